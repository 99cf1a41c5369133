/* Embeds the kernel source compiled into this library as the payload
 * returned by rainbow_shader_ptr/len (the counterpart of the reference's
 * include_str! of its WGSL, src/lib.rs:7). */
    .section .rodata
    .balign 16
    .global rainbow_kernel_source_begin
    .global rainbow_kernel_source_end
rainbow_kernel_source_begin:
    .incbin "rainbow_kernels.hip"
rainbow_kernel_source_end:
    .byte 0
    .section .note.GNU-stack,"",@progbits
