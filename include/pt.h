/*
 * libpt-compatible C API of the MI355X (gfx950) perturbative-triples
 * engine.  Drop-in for the reference header src/pt.h: same symbols, same
 * argument meaning, same array layouts, same return value.
 *
 * Layouts (reference src/pt.h:43-57, 79-93, 123-124, 156-157): all arrays
 * are row-major with the last index fastest.  Restricted calls take
 * [aaaa | abab] blocks, unrestricted calls [aaaa | abab | bbbb | baba];
 * the same-spin ovvv blocks are packed over b > c (index b(b-1)/2 + c).
 *
 * Differences in mechanism, not in meaning:
 *  - inputs are caller-owned host memory, copied to the GPU inside each
 *    call and not retained after it returns;
 *  - there is no MPI: one call computes the whole energy on one GPU
 *    (LIBPT_AMD_DEVICE selects which, default 0);
 *  - on failure (no GPU, allocation or HIP error) the library prints a
 *    message and exits with status 1, as the reference's err(1, ...) does.
 */
#ifndef LIBPT_PT_H
#define LIBPT_PT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Print the banner (reference: src/pt.h:26-27, src/misc.c:25-31). */
void libpt_print_banner(void);

/* Host scratch allocator hooks; NULL restores malloc/free
 * (reference: src/pt.h:29-35, src/misc.c:33-47). */
void libpt_set_malloc(void *(*fn)(size_t));
void libpt_set_free(void (*fn)(void *));

/* Restricted CCSD(T) correction, FP64 (reference: src/pt.h:37-66,
 * src/pt.c:454-479).  Sizes: d_ov, f_ov, t1: oa*va; t2, i_oovv:
 * 2*oa*oa*va*va; i_oovo: 2*oa*oa*va*oa; i_ovvv: oa*va*va*(va-1)/2 +
 * oa*va*va*va. */
double libpt_rpt(size_t oa, size_t va, const double *d_ov, const double *f_ov,
    const double *t1, const double *t2, const double *i_oovo,
    const double *i_oovv, const double *i_ovvv);

/* FP32 variant (reference: src/pt.h:68-71, src/ptmp.c:454-479):
 * contractions and per-element terms in FP32, energy summed in FP64. */
double libpt_rpt_mp(size_t oa, size_t va, const float *d_ov, const float *f_ov,
    const float *t1, const float *t2, const float *i_oovo,
    const float *i_oovv, const float *i_ovvv);

/* Unrestricted CCSD(T) correction (reference: src/pt.h:73-107,
 * src/pt.c:481-535). */
double libpt_upt(size_t oa, size_t va, size_t ob, size_t vb, const double *d_ov,
    const double *f_ov, const double *t1, const double *t2,
    const double *i_oovo, const double *i_oovv, const double *i_ovvv);

/* FP32 variant (reference: src/pt.h:109-112, src/ptmp.c:481-535). */
double libpt_upt_mp(size_t oa, size_t va, size_t ob, size_t vb,
    const float *d_ov, const float *f_ov, const float *t1, const float *t2,
    const float *i_oovo, const float *i_oovv, const float *i_ovvv);

/* Restricted EOM-CCSD(fT) correction (reference: src/pt.h:114-139,
 * src/pt.c:936-972).  An excitation energy omega enters through d_ov. */
double libpt_rft(size_t oa, size_t va, const double *d_ov, const double *f2_ov,
    const double *l1, const double *t2, const double *l2, const double *i_oovv,
    const double *i2_t2f2_oovo, const double *i3_ovvv, const double *i6_oovo,
    const double *i7_ovvv);

/* FP32 variant (reference: src/pt.h:141-145, src/ptmp.c:936-972). */
double libpt_rft_mp(size_t oa, size_t va, const float *d_ov, const float *f2_ov,
    const float *l1, const float *t2, const float *l2, const float *i_oovv,
    const float *i2_t2f2_oovo, const float *i3_ovvv, const float *i6_oovo,
    const float *i7_ovvv);

/* Unrestricted EOM-CCSD(fT) correction (reference: src/pt.h:147-172,
 * src/pt.c:974-1050). */
double libpt_uft(size_t oa, size_t va, size_t ob, size_t vb, const double *d_ov,
    const double *f2_ov, const double *l1, const double *t2, const double *l2,
    const double *i_oovv, const double *i2_t2f2_oovo, const double *i3_ovvv,
    const double *i6_oovo, const double *i7_ovvv);

/* FP32 variant (reference: src/pt.h:174-178, src/ptmp.c:974-1050). */
double libpt_uft_mp(size_t oa, size_t va, size_t ob, size_t vb,
    const float *d_ov, const float *f2_ov, const float *l1, const float *t2,
    const float *l2, const float *i_oovv, const float *i2_t2f2_oovo,
    const float *i3_ovvv, const float *i6_oovo, const float *i7_ovvv);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* LIBPT_PT_H */
