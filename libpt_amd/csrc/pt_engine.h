/*
 * Internal interface between the host engine and the device kernels of
 * libpt_amd (not part of the public C API).
 *
 * A spin case of one system is a list of *terms*.  Each term adds
 *     sign * sum_kappa S[s][kappa] * Q[kappa][u][w]
 * to the connected triples W[a][b][c] of one ijk item, where the single
 * index s and the pair (u, w) are the output indices a, b, c in some role
 * (role = which of a/b/c the S operand carries; (u, w) are the other two in
 * a, b, c order).  S and Q are strided slices of the input tensors whose
 * base offsets are linear in the occupied indices (i, j, k).
 *
 * The terms of every spin case of libpt (src/pt.c:236-290, :361-432,
 * :588-696, :773-915) are written in this form by pt_terms.cpp.
 */
#ifndef LIBPT_AMD_PT_ENGINE_H
#define LIBPT_AMD_PT_ENGINE_H

#include <stddef.h>
#include <stdint.h>

enum pt_form {
	PT_FORM_FP64 = 0,   /* FP64 operands, v_mfma_f64_16x16x4_f64 */
	PT_FORM_FP32 = 1,   /* FP32 operands, v_mfma_f32_16x16x4_f32 (_mp) */
	PT_FORM_BF16X3 = 2, /* hi+lo BF16 pairs, 3 products per kappa packed
	                       into v_mfma_f32_16x16x32_bf16 */
	PT_FORM_F16X3 = 3,  /* hi+lo FP16 pairs (scaled), the same with
	                       v_mfma_f32_16x16x32_f16 */
	PT_NFORM = 4
};

enum pt_kind { PT_AAA = 0, PT_AAB = 1 };
enum pt_role { PT_ROLE_A = 0, PT_ROLE_B = 1, PT_ROLE_C = 2 };

#define PT_MAX_TERMS 40
#define PT_MAX_BUFS 24

/* kappa per group of the packed Q operands: one group of one (u, w) pair
 * is 32 bytes: 4 FP64, 8 FP32, or 8 split (hi, lo) pairs stored as
 * hi 2q, hi 2q+1, lo 2q, lo 2q+1 for q = 0 .. 3 */
static inline int
pt_kc(int form)
{
	return form == 0 ? 4 : 8;
}

static inline int
pt_group_bytes(int form)
{
	(void)form;
	return 32;
}

/* groups per packed slab of K kappa (zero padded), rounded up to an even
 * count: a step over a 16-wide w range takes two groups */
static inline int64_t
pt_kgroups(int form, int64_t K)
{
	const int64_t kc = pt_kc(form);
	return 2 * ((K + 2 * kc - 1) / (2 * kc));
}

/* Workgroup geometry (the kernel's Geo): NW = 4 waves, tile 32 a x 32 b x
 * TC c with TC = 16 for FP64 (whose 16 x 16 x 16 accumulator block per
 * wave fills 128 registers) and 32 for FP32 and the split forms (FP32
 * accumulators), two workgroups per CU; or NW = 8 waves, tile 32 a x 32 b
 * x 32 c, one
 * workgroup per CU with four staging slots.  Each wave owns 16 a x 16 b x
 * CW c (CW = TC for four waves, 16 for eight).  (fT) forms its two W sets
 * one after the other in the same registers, so the geometry does not
 * depend on the W-set count. */
#define PT_TA 32
#define PT_TB 32

/* default wave count of a form: 4 (two workgroups per CU) measured ahead
 * of 8 at o=45, v=477 in every form; the tuner's record for a size, or
 * LIBPT_AMD_WAVES=4 or 8, overrides it for new plans */
static inline int
pt_nw_default(int form)
{
	(void)form;
	return 4;
}

static inline int
pt_tcw(int form, int nset, int nw)
{
	(void)nset;
	return nw == 8 || form != 0 ? 32 : 16;
}

static inline int
pt_cw(int form, int nset, int nw)
{
	return nw == 8 ? 16 : pt_tcw(form, nset, nw);
}

/*
 * S operands are read in the API layout (kappa contiguous).  Q operands
 * [slab][kappa][plane] are read from packed copies
 * [slab][kappa / KC][plane][kappa % KC] (zero padded in kappa), so the KC
 * values of one (u, w) pair form one contiguous group.
 */
struct pt_term {
	int64_t s_off[3];  /* S base offset (elements) = s_off . (i, j, k) */
	int64_t q_off[3];  /* Q slab offset: in the API layout (elements) when
	                      built, in bytes of the packed copy when launched */
	int64_t s_st;      /* stride of the single index s (kappa stride 1) */
	int64_t q_k;       /* plane size of Q (kappa stride in the API layout) */
	int64_t q_st[3];   /* strides of a, b, c in the Q plane (S role unused) */
	int32_t s_buf, q_buf;
	int32_t role, set, sign, K;
	int32_t packed;    /* Q pair (u > w) is a packed triangle u(u-1)/2 + w */
	int32_t ufast;     /* u is the contiguous Q index (else w) */
};

/* epilogue inputs (element type: double, or float for PT_FORM_FP32) */
struct pt_epi {
	const void *d1, *d2;  /* d_ov of the pair spin / odd spin */
	const void *u1, *u2;  /* t1 (T) or l1 (fT) */
	const void *f1, *f2;  /* f_ov (T) or f2_ov (fT) */
	const void *GA, *GB;  /* i_oovv same-spin / mixed blocks */
	const void *ZA, *ZB;  /* t2 (T) or l2 (fT) same-spin / mixed blocks */
};

struct pt_launch {
	const void *buf[PT_MAX_BUFS];
	const struct pt_term *terms; /* device array: per W set, role a, role b,
	                                role c */
	int32_t nterm1, nterm2;      /* per W set: roles a and b, role c */
	int32_t kind, nset;
	int64_t oa, va, ob, vb;      /* aaa: ob = oa, vb = va */
	int32_t na, nc;              /* extents of a (= b) and c */
	int32_t tcw;                 /* c tile width (pt_tcw) */
	int32_t nterm_a;             /* per W set: terms with role a */
	const int32_t *ijk;          /* [n_ijk][3] */
	const int32_t *tiles;        /* [n_tiles][3]: blocks (A, B, C), A >= B */
	int32_t n_tiles, nw;         /* tiles per ijk, waves per workgroup */
	int64_t item0, n_chunk;      /* items [item0, item0 + n_chunk) */
	double *partial;             /* energy of each item */
	double inv_scale[2];         /* per W set: 1 / (operand scale product) */
	void *wsave;                 /* two W sets: W of set 0 of each item
	                                of the launch (pt_wsave_bytes each,
	                                item item0 + n at slot n), else
	                                unused */
	int32_t nwsave;              /* save areas in wsave (>= n_chunk) */
	int32_t kgrp;                /* > 1: the launch covers whole ijk
	                                (item0 and n_chunk multiples of
	                                n_tiles) taken in groups of kgrp
	                                consecutive ijk, each group's items
	                                tile-major; 1: items in order */
	uint32_t *ctr;               /* non-null: a persistent grid of pgrid
	                                workgroups takes the items from these
	                                8 counters (one per XCD, zeroed by the
	                                launch), with one save area per
	                                workgroup (nwsave >= pgrid); null: one
	                                workgroup per item */
	int32_t pgrid;
	struct pt_epi epi;
};

#ifdef __cplusplus
extern "C" {
#endif
/* bytes of one item's saved W set: NW waves x CW blocks x 4 values x 64
 * lanes */
static inline int64_t
pt_wsave_bytes(int form, int nset, int nw)
{
	return (int64_t)nw * pt_cw(form, nset, nw) * 4 * 64 *
	    (form == 0 ? 8 : 4);
}

/* items per launch with two W sets: each item of a launch has its own save
 * area, so this bounds the areas to 4 GB (32768 items of 128 KB) */
static inline int64_t
pt_ft_chunk(int form, int nw)
{
	return ((int64_t)4 << 30) / pt_wsave_bytes(form, 2, nw);
}

/* kernel launchers (pt_kernel.hip); return hipError_t as int */
int pt_launch_tiles(int form, const struct pt_launch *p, void *stream);
/* out[s] = fixed-order sum of segment s's partials (as pt_launch_reduce) */
int pt_launch_reduce_batch(const struct pt_launch *segs,
    const int64_t *begin, int nseg, double *out, void *stream);
int pt_launch_reduce(const double *x, int64_t n, double *out, void *stream);
int pt_launch_segsum(const double *x, int64_t nseg, int64_t seglen,
    double *out, void *stream);
int pt_launch_split(const double *src, uint32_t *dst, int64_t n, int fp16,
    double scale, void *stream);
/* [R][K][PL] -> [R][ceil(K/KC)][PL][KC] (mode 0 FP64, 1 FP32) or
 * [R][ceil(K/KC)][PL][hi KC][lo KC] halves (mode 2 BF16, 3 FP16, of x*scale);
 * P1 > 0: the plane [P1][PL / P1] is stored transposed */
int pt_launch_pack(const void *src, void *dst, int64_t R, int64_t K,
    int64_t nkb, int64_t PL, int KC, int mode, double scale, int64_t P1,
    void *stream);
int pt_launch_absmax(const double *src, int64_t n, double *out, void *stream);
int pt_launch_copy(const double *src, double *dst, int64_t n, void *stream);
int pt_launch_synth(const void *spec, int tensor, void *dst, int64_t n,
    int fp32, void *stream);

/* host engine (pt_host.cpp) for pt_noalloc.hip: the device every call
 * runs on (made current; -1 on error), the calling thread's error text
 * (NULL clears it), and one call whose device buffers all come from
 * [scratch, scratch + cap) */
int pt_device_init(void);
void pt_set_error(const char *msg);
double pt_noalloc_energy(int call, int form, size_t oa, size_t va,
    size_t ob, size_t vb, const void *const *dev_arrays, void *scratch,
    size_t cap, void *stream);
#ifdef __cplusplus
}
#endif

#endif /* LIBPT_AMD_PT_ENGINE_H */
