/* rr_data.S -- published input tables linked into librocrand.so as read-only data.
 *   joe_kuo_20000.bin : Joe-Kuo (new-joe-kuo-6.21201) Sobol direction numbers,
 *                       dimensions 1..19999 (rocrand_amd/data/make_joe_kuo.py)
 *   mtgp32_11213.bin  : the 512 MTGP32 parameter sets for exponent 11213
 *                       (rocrand_amd/data/make_mtgp32_params.py)
 *   mtgp32_11213_fast.bin : the same sets as mtgp32_params_fast_t records, exported as
 *                       mtgp32dc_params_fast_11213 for the device API
 *                       (include/rocrand/rocrand_mtgp32_11213.h)
 *   mt19937_jumps.bin : MT19937 jump polynomials x^(i*2^1000), x^(256*i*2^1000) mod phi
 *                       (rocrand_amd/data/make_mt19937_jumps.c)
 * The build assembles this file with the data directory on the .incbin search path. */
    .section .note.GNU-stack,"",@progbits
    .section .rodata
    .p2align 8
    .globl rr_joe_kuo_begin
    .hidden rr_joe_kuo_begin
rr_joe_kuo_begin:
    .incbin "joe_kuo_20000.bin"
    .globl rr_joe_kuo_end
    .hidden rr_joe_kuo_end
rr_joe_kuo_end:
    .p2align 8
    .globl rr_mtgp32_params_begin
    .hidden rr_mtgp32_params_begin
rr_mtgp32_params_begin:
    .incbin "mtgp32_11213.bin"
    .globl rr_mtgp32_params_end
    .hidden rr_mtgp32_params_end
rr_mtgp32_params_end:
    .p2align 8
    .globl mtgp32dc_params_fast_11213
    .type mtgp32dc_params_fast_11213, @object
    .size mtgp32dc_params_fast_11213, 120832
mtgp32dc_params_fast_11213:
    .incbin "mtgp32_11213_fast.bin"
    .p2align 8
    .globl rr_mt19937_jumps_begin
    .hidden rr_mt19937_jumps_begin
rr_mt19937_jumps_begin:
    .incbin "mt19937_jumps.bin"
    .globl rr_mt19937_jumps_end
    .hidden rr_mt19937_jumps_end
rr_mt19937_jumps_end:
