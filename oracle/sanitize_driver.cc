// TEST INFRASTRUCTURE ONLY.  Runs the oracle's entry points on one particle
// file so that a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_oracle_sanitized.py) checks them for out-of-bounds accesses,
// use of uninitialised memory through the sanitizer's checks, signed
// overflow and similar undefined behaviour.  Prints the FNV-1a digest of
// every output so the sanitized build can be compared with the normal one.
//
//   sanitize_driver <particles.bin> <steps>
//
// The file holds N reference-layout particles (32 bytes each).

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
uint32_t oracle_build_plans(uint32_t n, uint8_t* slots);
uint32_t oracle_fnv1a(const void* data, size_t bytes);
void oracle_solve_fields(const uint8_t* slots, const void* particles, void* moments, void* fields, int compensated);
void oracle_step(const uint8_t* slots, void* particles, void* moments, void* fields, uint32_t steps, int compensated);
void oracle_solve_fields_f64(const uint8_t* slots, const void* particles, void* moments, double* fields64);
void oracle_accelerations(const void* prm, const void* particles, const void* moments, const void* fields,
                          float* out, int compensated);
void oracle_accelerations_f64(const void* prm, const void* particles, const void* moments, const double* fields64,
                              double* out);
void oracle_leaf_cells(const void* prm, const void* particles, const void* moments, uint32_t* out);
void oracle_direct_sum_f64(uint32_t n, const void* particles, double gravity, double softening, uint32_t ntargets,
                           const uint32_t* targets, double* out);
void oracle_conserved_f64(uint32_t n, const void* particles, double* out);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <particles.bin> <steps>\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (f == nullptr) return 2;
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (bytes <= 0 || bytes % 32 != 0) return 2;
  const uint32_t n = static_cast<uint32_t>(bytes / 32);
  std::vector<float> parts(8u * n);
  if (std::fread(parts.data(), 32, n, f) != n) return 2;
  std::fclose(f);
  const uint32_t steps = static_cast<uint32_t>(std::atoi(argv[2]));

  std::vector<uint8_t> slots(4096, 0);
  const uint32_t depth = oracle_build_plans(n, slots.data());
  uint32_t cells = 0;
  for (int k = 0; k < 4; ++k) cells |= static_cast<uint32_t>(slots[28 + k]) << (8 * k);
  const uint8_t* leaf_slot = slots.data() + 256u * depth;

  auto print = [](const char* what, const void* p, size_t b) { std::printf("%s %08x\n", what, oracle_fnv1a(p, b)); };
  for (int comp = 0; comp < 2; ++comp) {
    std::vector<int32_t> m(4u * cells, 0);
    std::vector<float> fl(12u * cells, 0.0f);
    oracle_solve_fields(slots.data(), parts.data(), m.data(), fl.data(), comp);
    print(comp ? "moments_c" : "moments", m.data(), 4u * m.size());
    print(comp ? "fields_c" : "fields", fl.data(), 4u * fl.size());
    std::vector<float> acc(4u * n);
    oracle_accelerations(leaf_slot, parts.data(), m.data(), fl.data(), acc.data(), comp);
    print(comp ? "acc_c" : "acc", acc.data(), 4u * acc.size());
    std::vector<float> p = parts;
    std::vector<int32_t> m2(4u * cells, 0);
    std::vector<float> f2(12u * cells, 0.0f);
    oracle_step(slots.data(), p.data(), m2.data(), f2.data(), steps, comp);
    print(comp ? "step_c" : "step", p.data(), 4u * p.size());
  }
  std::vector<int32_t> m(4u * cells, 0);
  std::vector<double> f64(12u * cells, 0.0);
  oracle_solve_fields_f64(slots.data(), parts.data(), m.data(), f64.data());
  print("fields64", f64.data(), 8u * f64.size());
  std::vector<double> acc64(4u * n);
  oracle_accelerations_f64(leaf_slot, parts.data(), m.data(), f64.data(), acc64.data());
  print("acc64", acc64.data(), 8u * acc64.size());
  std::vector<uint32_t> leaves(n);
  oracle_leaf_cells(leaf_slot, parts.data(), m.data(), leaves.data());
  print("leaves", leaves.data(), 4u * leaves.size());
  std::vector<uint32_t> targets;
  for (uint32_t t = 0; t < n && targets.size() < 16u; t += 1u + n / 16u) targets.push_back(t);
  std::vector<double> direct(4u * targets.size());
  oracle_direct_sum_f64(n, parts.data(), 1.0, 1e-3, static_cast<uint32_t>(targets.size()), targets.data(),
                        direct.data());
  print("direct", direct.data(), 8u * direct.size());
  double cons[8];
  oracle_conserved_f64(n, parts.data(), cons);
  print("conserved", cons, sizeof(cons));
  return 0;
}
