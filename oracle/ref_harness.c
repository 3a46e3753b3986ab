/*
 * ref_harness.c -- bulk entry points around the reference's own posit32 add and mul.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile (target `ref`) with the host
 * build of the reference's libRpack/kernel_entry.cl (p32_add :370, p32_mul :399), which
 * is compiled as OpenCL C for the host CPU.  Only the two arithmetic routines are
 * reachable; the GEMM kernels of that file are discarded at link time.
 */
#include <stddef.h>
#include <stdint.h>

typedef struct {
    uint32_t v;
} ref_p32; /* layout of posit32_t in kernel_entry.cl:82 */

ref_p32 p32_add(ref_p32, ref_p32);
ref_p32 p32_mul(ref_p32, ref_p32);

void ref_p32_add_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        ref_p32 x = {a[i]}, y = {b[i]};
        out[i] = p32_add(x, y).v;
    }
}

void ref_p32_mul_v(const uint32_t *a, const uint32_t *b, uint32_t *out, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        ref_p32 x = {a[i]}, y = {b[i]};
        out[i] = p32_mul(x, y).v;
    }
}
