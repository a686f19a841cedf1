/*
 * CDNA4 (gfx950) kernels of the perturbative-triples engine.
 *
 * k_tiles: one workgroup = one ijk item x one (A, B, C) tile of 32 a x 32 b
 * x TC c.  Its NW waves (4 or 8; Geo below) each own a 16 a x 16 b x CW c
 * block of the connected triples W in MFMA accumulators and never write W
 * to memory:
 *
 *   phase 1 -- accumulator layout [a][b]@c (one 16x16 MFMA block per c):
 *       terms whose single index is a  (A = S rows a,  B = Q[.][b][c])
 *       terms whose single index is b  (A = Q[.][a][c], B = S rows b)
 *   transpose through LDS to layout [a][c]@b (one block per b)
 *   phase 2 -- terms whose single index is c
 *       (A = Q[.][a][b], B = S rows c)
 *   epilogue -- for every (a, b, c) of the reference's triangle:
 *       disconnected term Z, denominator Dn, energy W (W - Z) / Dn
 *       ((fT): (W_L - Z_L) W_R / Dn), compensated FP64 sum per lane,
 *       one deterministic reduction per workgroup into partial[item];
 *       the Z and Dn operands that depend on b are staged in LDS once
 *       per item, those of a and c alone held in registers per lane.
 *
 * Operands go from HBM straight into LDS with buffer-load-to-LDS DMAs (Q
 * from the packed copies in 16-byte runs, S rows in dwords), several steps
 * ahead of the step computed; waves meet at fence-free barriers and wait
 * only for the DMAs of the step they read next.
 *
 * Forms: FP64 (v_mfma_f64_16x16x4_f64), FP32 (v_mfma_f32_16x16x4_f32),
 * split BF16 / FP16 (hi+lo pairs, hi*hi + hi*lo + lo*hi packed into the
 * k slots of v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulation, FP64
 * epilogue).
 *
 * MFMA lane layouts (checked on gfx950): A[m][k] lane m + 16*(k/KL),
 * B[k][n] lane n + 16*(k/KL) (KL = 1 for x4, 4 consecutive k for x16);
 * D[m][n] lane n + 16*(m/4) reg m%4 for f32 results, but
 * lane n + 16*(m%4) reg m/4 for v_mfma_f64_16x16x4_f64.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pt_engine.h"
#include "synth.h"

typedef double v2d __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

namespace {

template <int F> struct Form;

template <> struct Form<PT_FORM_FP64> {
	typedef double G;   /* operand element in HBM */
	typedef double L;   /* LDS element (one plane) */
	typedef double AE;  /* accumulator element */
	typedef v4d ACC;
	typedef double ER;  /* epilogue real */
	static constexpr int KC = 8;    /* kappa per type-A step */
	static constexpr bool ROW_F64 = true, SPLIT = false;
};

template <> struct Form<PT_FORM_FP32> {
	typedef float G;
	typedef float L;
	typedef float AE;
	typedef v4f ACC;
	typedef float ER;
	static constexpr int KC = 16;
	static constexpr bool ROW_F64 = false, SPLIT = false;
};

template <> struct Form<PT_FORM_BF16X3> {
	typedef uint32_t G;  /* hi in bits 0-15, lo in bits 16-31 */
	typedef uint32_t L;  /* LDS word: two hi or two lo halves */
	typedef float AE;
	typedef v4f ACC;
	typedef double ER;
	static constexpr int KC = 16;
	static constexpr bool ROW_F64 = false, SPLIT = true;
};

template <> struct Form<PT_FORM_F16X3> : Form<PT_FORM_BF16X3> {};

/* ---------------------------------------------------------------- MFMA */

__device__ __forceinline__ v4d
mma(double a, double b, v4d c)
{
	return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4f
mma(float a, float b, v4f c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

/* one K=32 bf16 / f16 MFMA (v_mfma_f32_16x16x32_{bf16,f16}, full rate on
 * gfx950) on operands of four 32-bit words (eight 16-bit values) per lane.
 * A and B fill their eight slots of each lane group with the same
 * (kappa, hi/lo) pattern, so each slot contributes one product of the
 * same kappa whatever the instruction's k order within the lane group. */
template <int F>
__device__ __forceinline__ v4f
mma32(v4u a, v4u b, v4f c)
{
	if constexpr (F == PT_FORM_BF16X3)
		return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
		    __builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0,
		    0, 0);
	else
		return __builtin_amdgcn_mfma_f32_16x16x32_f16(
		    __builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0,
		    0);
}

/* hi*hi + lo*hi + hi*lo (FP32 accumulation) of one split operand pair.
 * Operand words (LDS layout of the split forms): w0 = hi k0 k1,
 * w1 = lo k0 k1, w2 = hi k2 k3, w3 = lo k2 k3 (type A: four kappa per
 * lane group) or w0, w1 only (type B: two kappa).
 *   type A: slots (hi k0-3, lo k0-3) x (hi k0-3, hi k0-3) give hi*hi and
 *           lo*hi; slots (hi k0-3, 0) x (lo k0-3, 0) give hi*lo, first;
 *   type B: slots (hi k0 k1, lo k0 k1, hi k0 k1, 0 0) x
 *           (hi k0 k1, hi k0 k1, lo k0 k1, 0 0) give all three at once. */
template <int F, bool TA>
__device__ __forceinline__ v4f
mma_split(v4u a, v4u b, v4f c)
{
	if constexpr (TA) {
		c = mma32<F>(v4u{a[0], a[2], 0u, 0u}, v4u{b[1], b[3], 0u, 0u}, c);
		return mma32<F>(v4u{a[0], a[2], a[1], a[3]},
		    v4u{b[0], b[2], b[0], b[2]}, c);
	} else {
		return mma32<F>(v4u{a[0], a[1], a[0], 0u},
		    v4u{b[0], b[0], b[1], 0u}, c);
	}
}

/* ---------------------------------------------------------- epilogue */

__device__ __forceinline__ void
kahan(double &s, double &cmp, double x)
{
	double y = x - cmp;
	double t = s + y;
	cmp = (t - s) - y;
	s = t;
}

/*
 * Energy terms of one lane's NB elements (a, b0 + bb, c), bb = 0 .. NB-1,
 * in bb order, added to (esum, ecmp): the disconnected term Z, the
 * denominator Dn = d_ia + d_jb + d_kc and W (W - Z) / Dn, or for two W sets
 * (W_L - Z) W_R / Dn (src/pt.c:700,927).  The loads that depend on a, c or
 * (a, c) only are made once for the NB elements; those that depend on b
 * only are wave-uniform (b0 is); b is clamped into the arrays, so every
 * element's loads are made whether or not the triangle keeps it, and only
 * the sum is masked.  W values: ew[64 bb] (one set; the right set with two)
 * and ew[64 (bb + wl_ofs)] (the left set).  STC > 0: the workgroup has
 * staged in LDS the (b, c) values of i_oovv and t2, sbc[(32 n + bb) STC]
 * (sbc points at this lane's b0 and c) for n = G jk, Z jk, G ik, Z ik
 * (mixed spin: GB, ZB; same spin also G ji, Z ji), and the values that
 * depend on b only, sbv[32 n + bb] (sbv points at b0) for n = u i, f i,
 * u j, f j, (same spin: u k, f k,) d1 j; SBA (mixed spin): also the (b, a)
 * values of the same-spin blocks for ji, sba[(32 n + bb) 32] for n = GA,
 * ZA (sba points at b0 and this lane's a).
 */
template <int F, int KIND, int NSET, int NB, int STC, bool SBA>
__device__ __forceinline__ void
epi_block(const pt_launch &p, int i, int j, int k, int a, int c, int b0,
    const typename Form<F>::AE *ew, int wl_ofs,
    const typename Form<F>::ER *sbc, const typename Form<F>::ER *sba,
    const typename Form<F>::ER *sbv, double &esum, double &ecmp)
{
	typedef typename Form<F>::ER ER;
	const int na = p.na;
	const ER s0 = (ER)p.inv_scale[0], s1 = (ER)p.inv_scale[1];
	const ER *d1 = (const ER *)p.epi.d1;
	/* this element's term from its Z and Dn */
	auto term = [&](int bb, ER z, ER dn) {
		if constexpr (NSET == 1) {
			ER w = (ER)ew[64 * bb];
			if (Form<F>::SPLIT)
				w *= s0;
			return w * (w - z) / dn;
		} else {
			ER wl = (ER)ew[64 * (bb + wl_ofs)], wr = (ER)ew[64 * bb];
			if (Form<F>::SPLIT) {
				wl *= s0;
				wr *= s1;
			}
			return (wl - z) * wr / dn;
		}
	};
	if constexpr (KIND == PT_AAA) {
		/* src/pt.c:162-175, for (u, G) = (t1 or l1, i_oovv) and
		 * (f, Z) = (f_ov or f2_ov, t2 or l2) */
		const int64_t o = p.oa, v = p.va, vv = v * v;
		const ER *u = (const ER *)p.epi.u1, *f = (const ER *)p.epi.f1;
		const ER *G = (const ER *)p.epi.GA, *Z = (const ER *)p.epi.ZA;
		const int64_t jk = ((int64_t)j * o + k) * vv;
		const int64_t ik = ((int64_t)i * o + k) * vv;
		const int64_t ji = ((int64_t)j * o + i) * vv;
		const ER uia = u[i * v + a], uja = u[j * v + a], uka = u[k * v + a];
		const ER fia = f[i * v + a], fja = f[j * v + a], fka = f[k * v + a];
		const ER uic = u[i * v + c], ujc = u[j * v + c], ukc = u[k * v + c];
		const ER fic = f[i * v + c], fjc = f[j * v + c], fkc = f[k * v + c];
		const int64_t ac = (int64_t)a * v + c;
		const ER gjkac = G[jk + ac], gikac = G[ik + ac], gjiac = G[ji + ac];
		const ER zjkac = Z[jk + ac], zikac = Z[ik + ac], zjiac = Z[ji + ac];
		const ER dia = d1[i * v + a], dkc = d1[k * v + c];
#pragma unroll 1
		for (int bb = 0; bb < NB; ++bb) {
			const int b = b0 + bb;
			const int bq = b < na ? b : na - 1;
			const int64_t bc = (int64_t)bq * v + c;
			const int64_t ba = (int64_t)bq * v + a;
			ER uib, ujb, ukb, fib, fjb, fkb, djb;
			ER gjkbc, zjkbc, gikbc, zikbc, gjibc, zjibc;
			if constexpr (STC > 0) {
				uib = sbv[bb];
				fib = sbv[32 + bb];
				ujb = sbv[64 + bb];
				fjb = sbv[96 + bb];
				ukb = sbv[128 + bb];
				fkb = sbv[160 + bb];
				djb = sbv[192 + bb];
				gjkbc = sbc[bb * STC];
				zjkbc = sbc[(32 + bb) * STC];
				gikbc = sbc[(64 + bb) * STC];
				zikbc = sbc[(96 + bb) * STC];
				gjibc = sbc[(128 + bb) * STC];
				zjibc = sbc[(160 + bb) * STC];
			} else {
				uib = u[i * v + bq];
				fib = f[i * v + bq];
				ujb = u[j * v + bq];
				fjb = f[j * v + bq];
				ukb = u[k * v + bq];
				fkb = f[k * v + bq];
				djb = d1[j * v + bq];
				gjkbc = G[jk + bc];
				zjkbc = Z[jk + bc];
				gikbc = G[ik + bc];
				zikbc = Z[ik + bc];
				gjibc = G[ji + bc];
				zjibc = Z[ji + bc];
			}
			const ER zu = uia * gjkbc - uja * gikbc -
			    uka * gjibc - uib * gjkac + ujb * gikac +
			    ukb * gjiac - uic * G[jk + ba] + ujc * G[ik + ba] +
			    ukc * G[ji + ba];
			const ER zf = fia * zjkbc - fja * zikbc -
			    fka * zjibc - fib * zjkac + fjb * zikac +
			    fkb * zjiac - fic * Z[jk + ba] + fjc * Z[ik + ba] +
			    fkc * Z[ji + ba];
			const ER t = term(bb, zu + zf, dia + djb + dkc);
			if (a > b && b > c)
				kahan(esum, ecmp, (double)t);
		}
	} else {
		/* src/pt.c:177-185 and :433-442 */
		const int64_t oa = p.oa, va = p.va, ob = p.ob, vb = p.vb;
		const ER *u1 = (const ER *)p.epi.u1, *u2 = (const ER *)p.epi.u2;
		const ER *f1 = (const ER *)p.epi.f1, *f2 = (const ER *)p.epi.f2;
		const ER *GA = (const ER *)p.epi.GA, *GB = (const ER *)p.epi.GB;
		const ER *ZA = (const ER *)p.epi.ZA, *ZB = (const ER *)p.epi.ZB;
		const ER *d2 = (const ER *)p.epi.d2;
		const int64_t jk = ((int64_t)j * ob + k) * va;
		const int64_t ik = ((int64_t)i * ob + k) * va;
		const int64_t ji = ((int64_t)j * oa + i) * va;
		const ER uia = u1[i * va + a], fia = f1[i * va + a];
		const ER uja = u1[j * va + a], fja = f1[j * va + a];
		const ER u2c = u2[k * vb + c], f2c = f2[k * vb + c];
		const int64_t jkac = (jk + a) * vb + c, ikac = (ik + a) * vb + c;
		const ER gjkac = GB[jkac], zjkac = ZB[jkac];
		const ER gikac = GB[ikac], zikac = ZB[ikac];
		const ER dia = d1[i * va + a], dkc = d2[k * vb + c];
#pragma unroll 1
		for (int bb = 0; bb < NB; ++bb) {
			const int b = b0 + bb;
			const int bq = b < na ? b : na - 1;
			const int64_t jkbc = (jk + bq) * vb + c;
			const int64_t ikbc = (ik + bq) * vb + c;
			const int64_t jiba = (ji + bq) * va + a;
			ER uib, fib, ujb, fjb, djb;
			ER gjkbc, zjkbc, gikbc, zikbc, gjiba, zjiba;
			if constexpr (STC > 0) {
				uib = sbv[bb];
				fib = sbv[32 + bb];
				ujb = sbv[64 + bb];
				fjb = sbv[96 + bb];
				djb = sbv[128 + bb];
				gjkbc = sbc[bb * STC];
				zjkbc = sbc[(32 + bb) * STC];
				gikbc = sbc[(64 + bb) * STC];
				zikbc = sbc[(96 + bb) * STC];
			} else {
				uib = u1[i * va + bq];
				fib = f1[i * va + bq];
				ujb = u1[j * va + bq];
				fjb = f1[j * va + bq];
				djb = d1[j * va + bq];
				gjkbc = GB[jkbc];
				zjkbc = ZB[jkbc];
				gikbc = GB[ikbc];
				zikbc = ZB[ikbc];
			}
			if constexpr (SBA) {
				gjiba = sba[bb * 32];
				zjiba = sba[(32 + bb) * 32];
			} else {
				gjiba = GA[jiba];
				zjiba = ZA[jiba];
			}
			const ER z = (uia * gjkbc + fia * zjkbc)
			    - (uib * gjkac + fib * zjkac)
			    - (uja * gikbc + fja * zikbc)
			    + (ujb * gikac + fjb * zikac)
			    + (u2c * gjiba + f2c * zjiba);
			const ER t = term(bb, z, dia + djb + dkc);
			if (a > b)
				kahan(esum, ecmp, (double)t);
		}
	}
}

/* ------------------------------------------------- one staging step */

/* values the whole wave shares: readfirstlane tells the compiler so, which
 * keeps them (and everything derived from them) in scalar registers */
__device__ __forceinline__ int
uni(int x)
{
	return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ int64_t
uni64(int64_t x)
{
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
	const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
	return (int64_t)(((uint64_t)hi << 32) | lo);
}

/* Operand geometry of one term inside the workgroup tile (uniform), with
 * the byte bases of its S rows and of kappa group 0 of its packed Q slab;
 * a step (kappa chunk kc) then only moves both bases. */
struct Step {
	const char *sb0;     /* S element (s = 0, kappa = 0) of this ijk */
	const char *qb0;     /* packed Q group 0 of this ijk */
	int64_t sbytes;      /* S rows [0, sext) from sb0 */
	int64_t qgs;         /* bytes per kappa group of the packed slab */
	int64_t s_st, q_u, q_w;
	int role, set, K, packed;
	int s0, sn, sext, u0, uext, w0, wext;
};

/* T: the workgroup's LDS copy of a term; every field is read as a
 * wave-uniform value */
template <int ES, int GB>
__device__ __forceinline__ Step
make_step(const pt_launch &p, const pt_term &T, int i, int j, int k, int A0,
    int B0, int C0, int tcw)
{
	Step s;
	s.role = uni(T.role);
	s.set = uni(T.set);
	s.K = uni(T.K);
	s.packed = uni(T.packed);
	s.s_st = uni64(T.s_st);
	s.qgs = uni64(T.q_k) * GB;
	const int na = p.na, nc = p.nc;
	int ur, wr;
	if (s.role == PT_ROLE_A) {
		s.s0 = A0; s.sn = 32; s.sext = na;
		s.u0 = B0; s.uext = na; s.w0 = C0; s.wext = nc;
		ur = PT_ROLE_B; wr = PT_ROLE_C;
	} else if (s.role == PT_ROLE_B) {
		s.s0 = B0; s.sn = 32; s.sext = na;
		s.u0 = A0; s.uext = na; s.w0 = C0; s.wext = nc;
		ur = PT_ROLE_A; wr = PT_ROLE_C;
	} else {
		s.s0 = C0; s.sn = tcw; s.sext = nc;
		s.u0 = A0; s.uext = na; s.w0 = B0; s.wext = na;
		ur = PT_ROLE_A; wr = PT_ROLE_B;
	}
	s.q_u = uni64(T.q_st[ur]);
	s.q_w = uni64(T.q_st[wr]);
	const int64_t soff = uni64(T.s_off[0]) * i + uni64(T.s_off[1]) * j +
	    uni64(T.s_off[2]) * k;
	const int64_t qoff = uni64(T.q_off[0]) * i + uni64(T.q_off[1]) * j +
	    uni64(T.q_off[2]) * k;
	s.sb0 = (const char *)uni64((int64_t)p.buf[uni(T.s_buf)]) + soff * ES;
	s.qb0 = (const char *)uni64((int64_t)p.buf[uni(T.q_buf)]) + qoff;
	s.sbytes = (int64_t)s.sext * s.s_st * ES;
	return s;
}

/* buffer descriptors: loads at offsets at or beyond the descriptor's byte
 * count return zero, which handles tile edges, invalid pairs and the kappa
 * tail without branches */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t
brsrc(const void *base, int64_t bytes)
{
	const int nrec = bytes <= 0 ? 0 : bytes > 0x7fffffff ? 0x7fffffff :
	    (int)bytes;
	return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, nrec,
	    0x00020000);
}

#define PT_OOR 0x80000000u  /* an offset every descriptor rejects */

/* buffer store / load of one accumulator value at lane offset vofs plus a
 * constant offset sofs: the constant goes to a scalar register, so a run of
 * them costs no vector registers for addresses */
template <class AE>
__device__ __forceinline__ void
bstore_ae(__amdgpu_buffer_rsrc_t r, uint32_t vofs, int sofs, AE v)
{
	typedef uint32_t v2u __attribute__((ext_vector_type(2)));
	if constexpr (sizeof(AE) == 8)
		__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r,
		    (int)vofs, sofs, 0);
	else
		__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t,
		    v), r, (int)vofs, sofs, 0);
}

template <class AE>
__device__ __forceinline__ AE
bload_ae(__amdgpu_buffer_rsrc_t r, uint32_t vofs, int sofs)
{
	if constexpr (sizeof(AE) == 8)
		return __builtin_bit_cast(AE,
		    __builtin_amdgcn_raw_buffer_load_b64(r, (int)vofs, sofs, 0));
	else
		return __builtin_bit_cast(AE,
		    __builtin_amdgcn_raw_buffer_load_b32(r, (int)vofs, sofs, 0));
}


/*
 * Staging: NSLOT slots, each holding one step (term, kappa chunk) of the
 * tile, loaded NSLOT - 1 steps ahead of the step computed.
 * Slot (bytes):
 *   S: 32 rows of the chunk's kappa, packed at their data width (64 B for
 *      type A, 32 B for type B);
 *   Q: 32 u-rows of 1 KB of data + padding.  A u-row is the chunk of the
 *      tile's w range:
 *      type A (w range 16, roles a and b of a 16-wide c tile): 16 w x 64 B,
 *        i.e. two 32-byte packed kappa groups per (u, w);
 *      type B (w range 32: role c, and 32-wide c tiles): 32 w x 32 B, one
 *        packed group per (u, w).
 *      Either way a u-row is one 1 KB DMA run (lane l writes bytes 16 l ..
 *      16 l + 15).
 * MFMA lane group lg reads chunk lg of a type-A row (16 B: kappa 2lg, 2lg+1
 * for FP64, 4lg .. 4lg+3 for FP32) or element lg of a type-B row (8 B: kappa
 * lg for FP64, 2lg, 2lg+1 for FP32), the same mapping for both operands;
 * lane n of a lane group reads row n.  The Q row strides are those at
 * which such reads have no LDS bank conflicts on gfx950 (measured with
 * SQ_LDS_BANK_CONFLICT): 16-byte reads over 16 rows need a stride of 32
 * modulo 64 bytes (type A: 1056), 8-byte reads 16 modulo 64 (type B: 1040);
 * the other strides cost a 2-way conflict.  S rows are read by one
 * instruction per step, so their conflicts cost little, and packing them
 * at their data width takes one DMA instruction per wave and step fewer
 * (FP64 +1.4 % at o=45, v=477 against rows of 96 and 80 bytes).
 */
template <int NW, int NS>
struct Stg {
	static constexpr int S_ROW_A = 64, S_ROW_B = 32;
	static constexpr int Q_USTR_A = 1024 + 32, Q_USTR_B = 1024 + 16;
	static constexpr int S_BYTES = 2048;
	static constexpr int Q_BYTES = 32 * Q_USTR_A;
	static constexpr int SLOT = S_BYTES + Q_BYTES;
	static constexpr int NSLOT = NS;
	static constexpr int BYTES = NSLOT * SLOT;
	template <bool TA> static constexpr int s_row() {
		return TA ? S_ROW_A : S_ROW_B;
	}
	template <bool TA> static constexpr int q_ustr() {
		return TA ? Q_USTR_A : Q_USTR_B;
	}
};

/* lane offsets of the Q runs of one term: the 32 u-rows of a step are
 * dealt to the NW waves, QR = 32 / NW runs each (type A: w = lane / 4,
 * 16-byte chunk c = lane % 4 from group c / 2; type B: w = lane / 2, chunk
 * lane % 2); gstride = bytes per packed group of the slab */
template <bool TA, int NW>
__device__ __forceinline__ void
qtab(const Step &s, int wave, int lane, uint32_t *vo)
{
	const int w = TA ? (lane >> 2) : (lane >> 1);
	const uint32_t cofs = TA ? (uint32_t)((lane & 1) * 16 +
	    ((lane >> 1) & 1) * s.qgs) : (uint32_t)((lane & 1) * 16);
#pragma unroll
	for (int n = 0; n < 32 / NW; ++n) {
		const int ug = s.u0 + wave + NW * n, wg = s.w0 + w;
		const bool v = ug < s.uext && wg < s.wext && (!s.packed || ug > wg);
		const int64_t pidx = s.packed ? (int64_t)ug * (ug - 1) / 2 + wg :
		    (int64_t)ug * s.q_u + (int64_t)wg * s.q_w;
		vo[n] = v ? (uint32_t)(pidx * 32) + cofs : PT_OOR;
	}
}

/* DMA of step g (kappa groups g or 2g, 2g + 1) of a term into a slot */
template <bool TA, int NW>
__device__ __forceinline__ void
issue_q(const Step &s, int g, char *sQ, int wave, const uint32_t *vo)
{
	constexpr int NG = TA ? 2 : 1;
	const __amdgpu_buffer_rsrc_t rq = brsrc(s.qb0 + (int64_t)g * NG * s.qgs,
	    NG * s.qgs);
#pragma unroll
	for (int n = 0; n < 32 / NW; ++n)
		__builtin_amdgcn_raw_ptr_buffer_load_lds(rq,
		    (__attribute__((address_space(3))) void *)(sQ +
		    (wave + NW * n) * Stg<4, 2>::q_ustr<TA>()), 16, vo[n], 0, 0, 0);
}

/* wait until at most n vector-memory loads of this wave are outstanding */
template <int N>
__device__ __forceinline__ void
wait_vm()
{
	__asm__ volatile("s_waitcnt vmcnt(%0)" :: "n"(N) : "memory");
}

/*
 * S rows by DMA as well: the S region of a slot holds 32 rows at stride
 * S_ROW_A (64) or S_ROW_B (32) bytes, the step's chunk of kappa (in the API
 * element type).  The region is covered by NW x SD dword DMA instructions
 * per step: dword p = 64 (SD wave + d) + lane of the region is row
 * p / (stride / 4), dword p % (stride / 4) of it; rows outside the tile
 * fetch an offset every descriptor rejects (zero).  The term's sign is applied to the
 * accumulators (an exact negation before and after its steps), not to S.
 * Kappa beyond the term's extent read the next S elements, which meet
 * the zero padding of the packed Q operand.
 */
template <int NW>
struct SDma {
	static constexpr int SD = NW == 8 ? 1 : 2;   /* instructions per wave */
	static constexpr int BYTES = NW * SD * 256;  /* S region of a slot */
};

template <int F, int NW>
__device__ __forceinline__ void
stab(const Step &s, bool ta, int wave, int lane, uint32_t *so)
{
	constexpr int ES = (int)sizeof(typename Form<F>::G);
	constexpr int SD = SDma<NW>::SD;
	const int rd = (ta ? Stg<NW, 2>::S_ROW_A : Stg<NW, 2>::S_ROW_B) / 4;
	const int dd = (ta ? Form<F>::KC : Form<F>::KC / 2) * ES / 4;
#pragma unroll
	for (int d = 0; d < SD; ++d) {
		const int p = 64 * (SD * wave + d) + lane;
		const int r = p / rd, q = p - r * rd, sg = s.s0 + r;
		so[d] = (q < dd && r < s.sn && sg < s.sext) ?
		    (uint32_t)((int64_t)sg * s.s_st * ES + 4 * q) : PT_OOR;
	}
}

/* DMA of the S rows of the step at kappa kc into a slot's S region */
template <int F, int NW>
__device__ __forceinline__ void
issue_s(const Step &s, int kc, char *sS, int wave, const uint32_t *so)
{
	constexpr int ES = (int)sizeof(typename Form<F>::G);
	constexpr int SD = SDma<NW>::SD;
	const __amdgpu_buffer_rsrc_t rs = brsrc(s.sb0 + (int64_t)kc * ES,
	    s.sbytes - (int64_t)kc * ES);
#pragma unroll
	for (int d = 0; d < SD; ++d)
		__builtin_amdgcn_raw_ptr_buffer_load_lds(rs,
		    (__attribute__((address_space(3))) void *)(sS +
		    256 * (SD * wave + d)), 4, so[d], 0, 0, 0);
}

/* a split-form S operand: S rows hold API elements (hi | lo << 16 per
 * kappa); the MFMA operand words are hi k0 k1, lo k0 k1 (, hi k2 k3,
 * lo k2 k3), the layout of the packed Q groups */
__device__ __forceinline__ uint32_t
hi_pair(uint32_t e0, uint32_t e1)
{
	return (e0 & 0xffffu) | (e1 << 16);
}

__device__ __forceinline__ uint32_t
lo_pair(uint32_t e0, uint32_t e1)
{
	return (e0 >> 16) | (e1 & 0xffff0000u);
}

/* raw words (type A: four elements, type B: two, then zeros) -> the
 * operand words */
template <bool TA>
__device__ __forceinline__ v4u
split_words_s(v4u e)
{
	if constexpr (TA)
		return v4u{hi_pair(e[0], e[1]), lo_pair(e[0], e[1]),
		    hi_pair(e[2], e[3]), lo_pair(e[2], e[3])};
	else
		return v4u{hi_pair(e[0], e[1]), lo_pair(e[0], e[1]), 0u, 0u};
}

/*
 * Operand reads of one step from LDS, all of a wave's reads in one asm
 * block that ends by waiting for them (so the registers it returns hold
 * their data).  The outputs are early-clobber: a read's data may land
 * while later reads of the block still need the address registers.  Reading in asm keeps the compiler from placing a wait for
 * the DMAs in flight (into other slots) before these reads: the order of
 * DMAs and reads is kept by the explicit waits and barriers of the step
 * loop.  Type-B rows (32 B per w): S is 8 bytes at pf, Q value x is 8
 * bytes at pv + 32 x, returned in pairs (x = 2i in words 0-1, 2i + 1 in
 * words 2-3 of q[i]).  Type-A rows (64 B per w): 16 bytes at pf and at
 * pv + 64 x.
 */
#define PT_A(x) (unsigned)(size_t)(x)
#define PT_R2(i, o0, o1) "ds_read2_b64 %[q" #i "], %[pv] offset0:" #o0 \
	" offset1:" #o1 "\n\t"

__device__ __forceinline__ void
rd_b16(const char *pf, const char *pv, v2u &f, v4u (&q)[8])
{
	asm volatile("ds_read_b64 %[f], %[pf]\n\t"
	    PT_R2(0, 0, 4) PT_R2(1, 8, 12) PT_R2(2, 16, 20) PT_R2(3, 24, 28)
	    PT_R2(4, 32, 36) PT_R2(5, 40, 44) PT_R2(6, 48, 52) PT_R2(7, 56, 60)
	    "s_waitcnt lgkmcnt(0)"
	    : [f] "=&v"(f), [q0] "=&v"(q[0]), [q1] "=&v"(q[1]), [q2] "=&v"(q[2]),
	      [q3] "=&v"(q[3]), [q4] "=&v"(q[4]), [q5] "=&v"(q[5]),
	      [q6] "=&v"(q[6]), [q7] "=&v"(q[7])
	    : [pf] "v"(PT_A(pf)), [pv] "v"(PT_A(pv)) : "memory");
}

/* two S values (pf, pf + 512: rows 16 apart) and 16 Q values */
__device__ __forceinline__ void
rd_b16s2(const char *pf, const char *pv, v2u &f0, v2u &f1, v4u (&q)[8])
{
	asm volatile("ds_read_b64 %[f0], %[pf]\n\t"
	    "ds_read_b64 %[f1], %[pf] offset:512\n\t"
	    PT_R2(0, 0, 4) PT_R2(1, 8, 12) PT_R2(2, 16, 20) PT_R2(3, 24, 28)
	    PT_R2(4, 32, 36) PT_R2(5, 40, 44) PT_R2(6, 48, 52) PT_R2(7, 56, 60)
	    "s_waitcnt lgkmcnt(0)"
	    : [f0] "=&v"(f0), [f1] "=&v"(f1), [q0] "=&v"(q[0]), [q1] "=&v"(q[1]),
	      [q2] "=&v"(q[2]), [q3] "=&v"(q[3]), [q4] "=&v"(q[4]),
	      [q5] "=&v"(q[5]), [q6] "=&v"(q[6]), [q7] "=&v"(q[7])
	    : [pf] "v"(PT_A(pf)), [pv] "v"(PT_A(pv)) : "memory");
}

__device__ __forceinline__ void
rd_b32(const char *pf, const char *pv, v2u &f, v4u (&q)[16])
{
	asm volatile("ds_read_b64 %[f], %[pf]\n\t"
	    PT_R2(0, 0, 4) PT_R2(1, 8, 12) PT_R2(2, 16, 20) PT_R2(3, 24, 28)
	    PT_R2(4, 32, 36) PT_R2(5, 40, 44) PT_R2(6, 48, 52) PT_R2(7, 56, 60)
	    PT_R2(8, 64, 68) PT_R2(9, 72, 76) PT_R2(10, 80, 84) PT_R2(11, 88, 92)
	    PT_R2(12, 96, 100) PT_R2(13, 104, 108) PT_R2(14, 112, 116) PT_R2(15, 120, 124)
	    "s_waitcnt lgkmcnt(0)"
	    : [f] "=&v"(f), [q0] "=&v"(q[0]), [q1] "=&v"(q[1]), [q2] "=&v"(q[2]),
	      [q3] "=&v"(q[3]), [q4] "=&v"(q[4]), [q5] "=&v"(q[5]),
	      [q6] "=&v"(q[6]), [q7] "=&v"(q[7]), [q8] "=&v"(q[8]),
	      [q9] "=&v"(q[9]), [q10] "=&v"(q[10]), [q11] "=&v"(q[11]),
	      [q12] "=&v"(q[12]), [q13] "=&v"(q[13]), [q14] "=&v"(q[14]),
	      [q15] "=&v"(q[15])
	    : [pf] "v"(PT_A(pf)), [pv] "v"(PT_A(pv)) : "memory");
}

#define PT_R4(i, o) "ds_read_b128 %[q" #i "], %[pv] offset:" #o "\n\t"

__device__ __forceinline__ void
rd_a16(const char *pf, const char *pv, v4u &f, v4u (&q)[16])
{
	asm volatile("ds_read_b128 %[f], %[pf]\n\t"
	    PT_R4(0, 0) PT_R4(1, 64) PT_R4(2, 128) PT_R4(3, 192) PT_R4(4, 256) PT_R4(5, 320) PT_R4(6, 384)
	    PT_R4(7, 448) PT_R4(8, 512) PT_R4(9, 576) PT_R4(10, 640) PT_R4(11, 704) PT_R4(12, 768)
	    PT_R4(13, 832) PT_R4(14, 896) PT_R4(15, 960)
	    "s_waitcnt lgkmcnt(0)"
	    : [f] "=&v"(f), [q0] "=&v"(q[0]), [q1] "=&v"(q[1]), [q2] "=&v"(q[2]),
	      [q3] "=&v"(q[3]), [q4] "=&v"(q[4]), [q5] "=&v"(q[5]),
	      [q6] "=&v"(q[6]), [q7] "=&v"(q[7]), [q8] "=&v"(q[8]),
	      [q9] "=&v"(q[9]), [q10] "=&v"(q[10]), [q11] "=&v"(q[11]),
	      [q12] "=&v"(q[12]), [q13] "=&v"(q[13]), [q14] "=&v"(q[14]),
	      [q15] "=&v"(q[15])
	    : [pf] "v"(PT_A(pf)), [pv] "v"(PT_A(pv)) : "memory");
}

/* keep everything before it (the operand reads and the DMA issue) ahead
 * of everything after it (the MFMAs), and the MFMAs ahead of the wait */
__device__ __forceinline__ void
sched_fence()
{
	__builtin_amdgcn_sched_barrier(0);
}

/* operand x of a type-B read pair array: 8 bytes as 2 words */
__device__ __forceinline__ v2u
pair_of(const v4u *q, int x)
{
	const v4u t = q[x >> 1];
	return (x & 1) ? v2u{t[2], t[3]} : v2u{t[0], t[1]};
}

/* MFMAs of one type-B operand pair (8 bytes each) into c; SA: the first
 * operand is the S value (role a), else the second */
template <int F>
__device__ __forceinline__ typename Form<F>::ACC
mma_b(v2u a, v2u b, typename Form<F>::ACC c)
{
	if constexpr (F == PT_FORM_FP64) {
		return mma(__builtin_bit_cast(double, a),
		    __builtin_bit_cast(double, b), c);
	} else if constexpr (F == PT_FORM_FP32) {
		/* whole vectors are bit-cast, then indexed: a bit cast of one
		 * element (bit_cast(float, a[1])) reads element 0 with this
		 * compiler */
		const v2f x = __builtin_bit_cast(v2f, a), y = __builtin_bit_cast(v2f,
		    b);
		c = mma(x[0], y[0], c);
		return mma(x[1], y[1], c);
	} else {
		return mma_split<F, false>(v4u{a[0], a[1], 0u, 0u},
		    v4u{b[0], b[1], 0u, 0u}, c);
	}
}

/* the same for type-A operands (16 bytes each) */
template <int F>
__device__ __forceinline__ typename Form<F>::ACC
mma_a(v4u a, v4u b, typename Form<F>::ACC c)
{
	if constexpr (F == PT_FORM_FP64) {
		const v2d x = __builtin_bit_cast(v2d, a), y = __builtin_bit_cast(v2d,
		    b);
		c = mma(x[0], y[0], c);
		return mma(x[1], y[1], c);
	} else if constexpr (F == PT_FORM_FP32) {
		const v4f x = __builtin_bit_cast(v4f, a), y = __builtin_bit_cast(v4f,
		    b);
#pragma unroll
		for (int e = 0; e < 4; ++e)
			c = mma(x[e], y[e], c);
		return c;
	} else {
		return mma_split<F, true>(a, b, c);
	}
}

/* an S value for the MFMA: the split forms regroup the halves of the API
 * elements into hi and lo pair words */
template <int F, bool TA>
__device__ __forceinline__ v4u
s_operand(v4u raw)
{
	if constexpr (Form<F>::SPLIT)
		return split_words_s<TA>(raw);
	else
		return raw;
}

/*
 * Phase 1 (accumulator block x = c offset, D[a][b]) over one staged step:
 *   role a (RA): A = S[a] (fixed), B = Q[b][c = wc + x];
 *   role b:      A = Q[a][c = wc + x],  B = S[b] (fixed).
 * TA: type-A rows (64 B per w), else type-B rows (32 B per w); wc: the
 * wave's first c within the tile (type B only).  The step's operands are
 * read first; mid() (the DMAs of a later step) is issued after the reads
 * and before the MFMAs.
 */
template <int F, int TCW, bool RA, bool TA, class Mid>
__device__ __forceinline__ void
mma_p1(typename Form<F>::ACC (&ac)[TCW], const char *sS, const char *sQ,
    int wa, int wb, int wc, int ln, int lg, Mid mid)
{
	constexpr int NB = TA ? 16 : 8, WB = TA ? 64 : 32;
	const int ws = RA ? wa : wb, wu = RA ? wb : wa;
	const char *pf = sS + (16 * ws + ln) * Stg<4, 2>::s_row<TA>() + NB * lg;
	const char *pv = sQ + (16 * wu + ln) * Stg<4, 2>::q_ustr<TA>() + NB * lg +
	    WB * wc;
	if constexpr (TA) {
		static_assert(TCW == 16, "type-A steps have 16 w");
		v4u fr, q[16];
		rd_a16(pf, pv, fr, q);
		mid();
		sched_fence();
		const v4u f = s_operand<F, true>(fr);
#pragma unroll
		for (int x = 0; x < 16; ++x)
			ac[x] = RA ? mma_a<F>(f, q[x], ac[x]) : mma_a<F>(q[x], f, ac[x]);
	} else {
		v2u fr;
		v4u q[TCW / 2];
		if constexpr (TCW == 16)
			rd_b16(pf, pv, fr, q);
		else
			rd_b32(pf, pv, fr, q);
		mid();
		sched_fence();
		const v4u fw = s_operand<F, false>(v4u{fr[0], fr[1], 0u, 0u});
		const v2u f = v2u{fw[0], fw[1]};
#pragma unroll
		for (int x = 0; x < TCW; ++x)
			ac[x] = RA ? mma_b<F>(f, pair_of(q, x), ac[x]) :
			    mma_b<F>(pair_of(q, x), f, ac[x]);
	}
	sched_fence();
}

/*
 * Phase 2 (block x = 16 cb + b, D[a][c]): role c: A = Q[a][b], B = S[c]
 * (rows wc + 16 cb + ln); always type-B rows; reads, mid(), MFMAs as in
 * phase 1.
 */
template <int F, int TCW, class Mid>
__device__ __forceinline__ void
mma_p2(typename Form<F>::ACC (&ac)[TCW], const char *sS, const char *sQ,
    int wa, int wb, int wc, int ln, int lg, Mid mid)
{
	constexpr int NCB = TCW / 16;
	static_assert(NCB == 1 || NCB == 2, "16 or 32 c per wave");
	const char *pv = sQ + (16 * wa + ln) * Stg<4, 2>::Q_USTR_B + 16 * wb * 32 +
	    8 * lg;
	const char *pf = sS + (wc + ln) * Stg<4, 2>::S_ROW_B + 8 * lg;
	v2u fr[2];
	v4u q[8];
	if constexpr (NCB == 1)
		rd_b16(pf, pv, fr[0], q);
	else
		rd_b16s2(pf, pv, fr[0], fr[1], q);
	mid();
	sched_fence();
#pragma unroll
	for (int cb = 0; cb < NCB; ++cb) {
		const v4u fw = s_operand<F, false>(v4u{fr[cb][0], fr[cb][1], 0u,
		    0u});
		const v2u f = v2u{fw[0], fw[1]};
#pragma unroll
		for (int b = 0; b < 16; ++b)
			ac[16 * cb + b] = mma_b<F>(pair_of(q, b), f, ac[16 * cb + b]);
	}
	sched_fence();
}

/* -------------------------------------------------------------- kernel */

/*
 * Workgroup geometry, per wave count NW (pt_tcw, pt_cw in pt_engine.h give
 * the same numbers to the host):
 *   NW = 4: tile 32 a x 32 b x TC c, TC = 32 for FP32 and the split forms
 *           (whose 32 FP32 accumulator blocks of 4 registers fit beside a
 *           step's operands), 16 for FP64 (16 blocks of 8 registers); each
 *           wave owns 16 a x 16 b x TC c; two slots, two workgroups per CU,
 *           so one's staging and barriers overlap the other's MFMAs.  With
 *           TC = 32 every step is type B and the role-c operand is reused
 *           by 32 c (split-BF16 at o=45, v=477: 130 against 111 TFLOP/s
 *           with TC = 16).
 *   NW = 8: tile 32 a x 32 b x 32 c, each wave 16 a x 16 b x 16 c (wave =
 *           wa + 2 wb + 4 wc); one workgroup per CU with four slots (three
 *           steps in flight).  Every step is type B (32 w per u-row); the
 *           role-c operand is reused by 32 c instead of 16, and the CU holds
 *           the same accumulators as two 4-wave workgroups.
 */
template <int F, int NSET, int NW>
struct Geo {
	static constexpr int TC = NW == 8 || F != PT_FORM_FP64 ? 32 : 16;
	static constexpr int CW = NW == 8 ? 16 : TC;   /* c per wave */
	static constexpr int THREADS = 64 * NW;
	static constexpr int NS = NW == 8 ? 4 : 2;     /* staging slots */
	static constexpr int WG_PER_CU = NW == 8 ? 1 : 2;
	static constexpr int QR = 32 / NW;             /* Q runs per wave */
};

/*
 * One work item (ijk, tile) of a spin case: everything between reading the
 * item and writing its energy to partial[it].  KIND is the spin case,
 * PT_AAA or PT_AAB (only the epilogue depends on it).  With two W sets the
 * sets are formed one after the other; wsv is the item's save area for W
 * of set 0 (4 CW values per lane per wave).
 */
template <int F, int KIND, int NSET, int NW>
__device__ __forceinline__ void
tile_item(const pt_launch &p, int64_t it, typename Form<F>::AE *wsv)
{
	typedef Form<F> FM;
	typedef Geo<F, NSET, NW> GM;
	typedef Stg<NW, GM::NS> SG;
	typedef typename FM::G G;
	typedef typename FM::AE AE;
	typedef typename FM::ACC ACC;
	constexpr int KC = FM::KC;
	constexpr int TC = GM::TC, CW = GM::CW;
	constexpr int NCB = CW / 16;
	constexpr int NT = GM::THREADS;
	constexpr int GB = 32;  /* bytes per packed kappa group and pair */
	constexpr int SMEM = SG::BYTES;
	static_assert(SMEM >= NW * 2 * 8 * 64 * (int)sizeof(AE) &&
	    SMEM >= NW * 16 * 64 * (int)sizeof(AE),
	    "epilogue slots exceed the stage area");
	static_assert(SG::SLOT >= 4 * 4 * 16 * 17 * (int)sizeof(AE),
	    "transpose round exceeds a slot");
	static_assert(NSET == 1 || NSET == 2, "one or two W sets");
	static_assert(SDma<NW>::BYTES <= SG::S_BYTES,
	    "the S DMAs of a step stay inside the slot's S region");

	__shared__ __attribute__((aligned(16))) char smem[SMEM];
	__shared__ double red[NW];
	__shared__ pt_term sterm[PT_MAX_TERMS];
	AE *sX = (AE *)smem;

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wa = wave & 1, wb = (wave >> 1) & 1, wcx = wave >> 2;
	const int wc = 16 * wcx;       /* the wave's first c in the tile */
	const int ln = lane & 15, lg = lane >> 4;

	const bool same = KIND == PT_AAA;
	const int64_t ijk_id = it / p.n_tiles;
	const int tile = (int)(it - ijk_id * p.n_tiles);
	const int i = p.ijk[3 * ijk_id], j = p.ijk[3 * ijk_id + 1];
	const int k = p.ijk[3 * ijk_id + 2];
	const int A0 = p.tiles[3 * tile] * 32, B0 = p.tiles[3 * tile + 1] * 32;
	const int C0 = p.tiles[3 * tile + 2] * TC;
	const int na = p.na, nc = p.nc;
	const int a_lo = A0 + 16 * wa, b_lo = B0 + 16 * wb, c_lo = C0 + wc;
	/* does this wave's block hold any element of the triangle? */
	const bool act = a_lo < na && b_lo < na && a_lo + 15 > b_lo &&
	    c_lo < nc && (!same || b_lo + 15 > c_lo);

	ACC acc[CW];
#pragma unroll
	for (int x = 0; x < CW; ++x)
		acc[x] = ACC{0, 0, 0, 0};
	/* save area of W set 0: value (r, x) of (wave, lane) at element
	 * ((r CW + x) NW + wave) 64 + lane */
	constexpr int WST = NW * 64 * (int)sizeof(AE);
	const __amdgpu_buffer_rsrc_t wr = brsrc(wsv, NSET == 2 ?
	    (int64_t)4 * CW * WST : 0);
	const uint32_t wvo = (uint32_t)((wave * 64 + lane) * sizeof(AE));

	/* the term table, copied once into LDS: a term change then reads it
	 * with LDS loads instead of waiting on global memory.  Terms of one W
	 * set: role a, role b, role c; set 1 follows set 0. */
	const int nset_terms = uni(p.nterm1 + p.nterm2);
	const int nterm = NSET * nset_terms;
	{
		const uint32_t *src = (const uint32_t *)p.terms;
		uint32_t *dst = (uint32_t *)sterm;
		const int nw = nterm * (int)(sizeof(pt_term) / 4);
		for (int w = tid; w < nw; w += NT)
			dst[w] = src[w];
		__syncthreads();
	}
	constexpr int ES = (int)sizeof(G);
	auto step_of = [&](int tn) {
		return make_step<ES, GB>(p, sterm[tn], i, j, k, A0, B0, C0, TC);
	};

	/* [a][b]@c -> [a][c]@b through LDS scratch, per register r; with eight
	 * waves in two rounds of four (one slot holds four waves' scratch).
	 * Every wave passes every barrier. */
	auto transpose = [&](AE *scratch) {
		constexpr int NR = NW / 4;
#pragma unroll
		for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
		for (int r = 0; r < 4; ++r)
#pragma unroll
		for (int g = 0; g < NR; ++g) {
			const bool mine = (wave >> 2) == g;
			AE *X = scratch + (wave & 3) * (4 * 16 * 17);
			__syncthreads();
			if (mine) {
#pragma unroll
				for (int c = 0; c < 16; ++c)
					X[(lg * 16 + ln) * 17 + c] = acc[16 * cb + c][r];
			}
			__syncthreads();
			if (mine) {
#pragma unroll
				for (int b = 0; b < 16; ++b)
					acc[16 * cb + b][r] = X[(lg * 16 + b) * 17 + ln];
			}
		}
		__syncthreads();
	};

	{
		/*
		 * Steps n = 0 .. N-1 walk the terms (role a, role b, role c) and
		 * their kappa chunks; step n lives in slot n % NSLOT.  Iteration n
		 * issues the DMAs (Q and S) of step n + D (D = NSLOT - 1) into the
		 * slot step n - 1 used, computes step n, waits until this wave's
		 * DMAs of step n + 1 have landed and meets the other waves at a
		 * barrier without a memory fence, so that the DMAs of steps n + 2
		 * .. n + D stay in flight across it.
		 */
		constexpr int NS = SG::NSLOT, D = NS - 1;
		constexpr int KSA = KC, KSB = KC / 2;   /* kappa per step, type A / B */
		constexpr int VMS = GM::QR + SDma<NW>::SD; /* DMAs per step per wave */
		const int wv = __builtin_amdgcn_readfirstlane(wave);
		const int nta = uni(p.nterm_a), nt1 = uni(p.nterm1);
		/* type A: a 16-wide w range (roles a and b of a 16-wide c tile) */
		auto type_a = [&](int role) { return TC == 16 && role != PT_ROLE_C; };

		/* prefetch walker: term pt, chunk pk of step pn (the next to issue) */
		int pt = 0, pk = 0, pn = 0;
		Step pre = step_of(0);
		bool pa = type_a(pre.role);
		uint32_t vo[GM::QR], so[SDma<NW>::SD];
		auto term_tables = [&]() {
			if (pa)
				qtab<true, NW>(pre, wv, lane, vo);
			else
				qtab<false, NW>(pre, wv, lane, vo);
			stab<F, NW>(pre, pa, wv, lane, so);
		};
		term_tables();
		/* number of steps */
		int N = 0;
		for (int tt = 0; tt < nterm; ++tt) {
			const int kk = uni(sterm[tt].K);
			const int ks = type_a(uni(sterm[tt].role)) ? KSA : KSB;
			N += (kk + ks - 1) / ks;
		}
		/* issue step pn into its slot */
		auto issue = [&](char *slot) {
			if (pa)
				issue_q<true, NW>(pre, pk / KSA, slot + SG::S_BYTES, wv, vo);
			else
				issue_q<false, NW>(pre, pk / KSB, slot + SG::S_BYTES, wv,
				    vo);
			issue_s<F, NW>(pre, pk, slot, wv, so);
			++pn;
			pk += pa ? KSA : KSB;
			if (pk >= pre.K) {
				pk = 0;
				if (++pt < nterm) {
					pre = step_of(pt);
					pa = type_a(pre.role);
					term_tables();
				}
			}
		};
		/* wait until this wave's DMAs of step m (issued) have landed: at
		 * most (pn - 1 - m) later steps' DMAs may stay outstanding */
		auto wait_step = [&](int m) {
			const int later = pn - 1 - m;
			if (D >= 3 && later >= 2)
				wait_vm<2 * VMS>();
			else if (D >= 2 && later >= 1)
				wait_vm<VMS>();
			else
				wait_vm<0>();
		};
		/* the other waves' DMAs of that step: a barrier that waits on no
		 * memory operation (the compiler's barrier with a fence would wait
		 * for every DMA in flight) */
		auto step_barrier = [&]() {
			__asm__ volatile("" ::: "memory");
			__builtin_amdgcn_s_barrier();
			__asm__ volatile("" ::: "memory");
		};

		/* prologue: steps 0 .. D-1 in flight, step 0 landed */
		for (int m = 0; m < D; ++m)
			if (pn < N)
				issue(smem + (pn % NS) * SG::SLOT);
		wait_step(0);
		step_barrier();

		/* compute walker: term ct, chunk ck of step n; the accumulators
		 * hold W negated while cneg is set (the sign of the terms being
		 * added) */
		int ct = 0, ck = 0, cK = uni(sterm[0].K);
		bool cneg = false;
		int n = 0;
		auto set_sign = [&](bool neg) {
			if (neg != cneg) {
#pragma unroll
				for (int x = 0; x < CW; ++x)
					acc[x] = -acc[x];
				cneg = neg;
			}
		};
		/* steps of phase PH up to term tend (exclusive) */
		auto run_phase = [&](auto ph, int tend) {
			constexpr int PH = decltype(ph)::value;
			constexpr bool TA = PH < 2 && TC == 16;
			while (ct < tend) {
				set_sign(uni(sterm[ct].sign) < 0);
				/* step n reads slot rsl; after its operand reads, the
				 * DMAs of step n + D go to dsl, the slot step n - 1 used.
				 * The two never overlap within a step, which the
				 * restrict qualifiers tell the compiler, so that it
				 * places no wait for those DMAs before the reads; the
				 * order across steps comes from wait_step and the
				 * barrier. */
				auto body = [&](const char *__restrict__ rsl,
				    char *__restrict__ dsl) {
					auto mid = [&]() {
						if (pn < N)
							issue(dsl);
					};
					if constexpr (PH < 2)
						mma_p1<F, CW, PH == 0, TA>(acc, rsl,
						    rsl + SG::S_BYTES, wa, wb, wc, ln, lg, mid);
					else
						mma_p2<F, CW>(acc, rsl, rsl + SG::S_BYTES, wa, wb,
						    wc, ln, lg, mid);
				};
				body(smem + (n % NS) * SG::SLOT,
				    smem + (pn % NS) * SG::SLOT);
				if (n + 1 < N)
					wait_step(n + 1);
				step_barrier();
				++n;
				ck += TA ? KSA : KSB;
				if (ck >= cK) {
					ck = 0;
					if (++ct < nterm)
						cK = uni(sterm[ct].K);
				}
			}
			set_sign(false);
		};
#pragma unroll 1
		for (int s = 0; s < NSET; ++s) {
			const int t0 = s * nset_terms;
			run_phase(std::integral_constant<int, 0>(), t0 + nta);
			run_phase(std::integral_constant<int, 1>(), t0 + nt1);
			/* the slot of the last phase-1 step is free until step
			 * n + D issues into it: transpose through it */
			transpose((AE *)(smem + ((n + NS - 1) % NS) * SG::SLOT));
			run_phase(std::integral_constant<int, 2>(), t0 + nset_terms);
			if (NSET == 2 && s == 0) {
				/* W of set 0 to the save area, lane-contiguous per
				 * (r, x, wave); the accumulators restart for set 1 */
#pragma unroll
				for (int r = 0; r < 4; ++r)
#pragma unroll
					for (int x = 0; x < CW; ++x) {
						bstore_ae<AE>(wr, wvo, (r * CW + x) * WST,
						    acc[x][r]);
						acc[x][r] = 0;
					}
				/* the waits of later steps count outstanding
				 * memory operations; let none of them be a store */
				wait_vm<0>();
			}
		}
	}

	/* epilogue: layout [a][c]@b, block x = 16 cb + b; lane holds
	 * c = C0 + wc + 16 cb + ln.  Each register column r of each 16-wide
	 * c block goes through this lane's own LDS slots (BH b per W set at a
	 * time) so the element loop runs with a runtime index. */
	double esum = 0.0, ecmp = 0.0;
	const int b0u = uni(b_lo);  /* the wave's first b */
	/* the tile's epilogue operands that do not depend on a or c alone,
	 * staged once in LDS behind the W values where they fit: the (b, c)
	 * blocks of i_oovv and t2, [G jk, Z jk, G ik, Z ik (, G ji, Z ji)]
	 * [32 b][TC c] (mixed spin: the GB, ZB blocks; same spin: GA, ZA with
	 * ji), the b vectors [u i, f i, u j, f j, (u k, f k,) d1 j][32 b] and,
	 * for the mixed spin case where LDS allows, the (b, a) blocks of the
	 * same-spin ones for ji, [GA, ZA][32 b][32 a] (rows and columns beyond
	 * the arrays repeat the last one; the epilogue masks those elements).
	 * The previous step's barrier ended every read of the slots, and the
	 * barrier of the first W column below orders these writes before the
	 * reads. */
	typedef typename FM::ER ER;
	constexpr int NPB = same ? 6 : 4;     /* (b, c) blocks */
	constexpr int NBC = NPB * 32 * TC, NBV = (NPB + 1) * 32;
	constexpr int NBA = 2 * 32 * 32;
	constexpr int ES_ = (int)sizeof(ER);
	/* W passes of BH b each: 8 where W of 16 b (both sets) would leave no
	 * room for the staged operands */
	constexpr int BH = NW * NSET * 16 * 64 * (int)sizeof(AE) +
	    (NBC + NBV + (same ? 0 : NBA)) * ES_ <= SMEM ? 16 : 8;
	constexpr int EBYTES = NW * NSET * BH * 64 * (int)sizeof(AE);
	AE *E = sX + wave * (NSET * BH * 64) + lane;
	constexpr bool STG = EBYTES + (NBC + NBV) * ES_ <= SMEM &&
	    NBC % NT == 0 && NBV <= NT;
	constexpr bool SBA = !same && STG && NBA % NT == 0 &&
	    EBYTES + (NBC + NBV + NBA) * ES_ <= SMEM;
	ER *const stg = (ER *)(smem + EBYTES);
	const ER *sbc = stg + 16 * wb * TC + wc + ln;
	const ER *sbv = stg + NBC + 16 * wb;
	const ER *sba = stg + NBC + NBV + 16 * wb * 32 + 16 * wa;
	if constexpr (STG) {
		/* block n: (G, Z)[n & 1] of pair n / 2 = jk, ik (, ji); mixed
		 * spin: rows (jk + b) vb, same spin: jk + b v, v = vb */
		const ER *G = (const ER *)(same ? p.epi.GA : p.epi.GB);
		const ER *Z = (const ER *)(same ? p.epi.ZA : p.epi.ZB);
		const int64_t po = same ? p.oa : p.ob;
		const int64_t pr = same ? p.va * p.va : p.va;
		const int64_t pjk = ((int64_t)j * po + k) * pr;
		const int64_t pik = ((int64_t)i * po + k) * pr;
		const int64_t pji = ((int64_t)j * p.oa + i) * pr;
#pragma unroll
		for (int q = 0; q < NBC / NT; ++q) {
			const int n = tid + q * NT;
			const int arr = n / (32 * TC), bl = (n / TC) % 32, cl = n % TC;
			const int b = B0 + bl < na ? B0 + bl : na - 1;
			const int c = C0 + cl < nc ? C0 + cl : nc - 1;
			const int64_t pb = arr < 2 ? pjk : arr < 4 ? pik : pji;
			const int64_t row = same ? pb + (int64_t)b * p.vb :
			    (pb + b) * p.vb;
			stg[n] = ((arr & 1) ? Z : G)[row + c];
		}
		/* vector n: (u, f)[n & 1] at occupied index (i, j, k)[n / 2], the
		 * last one d1 at j */
		if (tid < NBV) {
			const int arr = tid / 32, bl = tid % 32;
			const int b = B0 + bl < na ? B0 + bl : na - 1;
			const int occ = arr < 2 ? i : arr < 4 || arr == NPB ? j : k;
			const ER *src = arr == NPB ? (const ER *)p.epi.d1 :
			    (arr & 1) ? (const ER *)p.epi.f1 : (const ER *)p.epi.u1;
			stg[NBC + tid] = src[occ * p.va + b];
		}
	}
	if constexpr (SBA) {
		const ER *GA = (const ER *)p.epi.GA, *ZA = (const ER *)p.epi.ZA;
		const int64_t ji = ((int64_t)j * p.oa + i) * p.va;
#pragma unroll
		for (int q = 0; q < NBA / NT; ++q) {
			const int n = tid + q * NT;
			const int arr = n / 1024, bl = (n / 32) % 32, al = n % 32;
			const int b = B0 + bl < na ? B0 + bl : na - 1;
			const int a = A0 + al < na ? A0 + al : na - 1;
			stg[NBC + NBV + n] = (arr ? ZA : GA)[(ji + b) * p.va + a];
		}
	}
#pragma unroll
	for (int r = 0; r < 4; ++r)
#pragma unroll
	for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
	for (int h = 0; h < 16 / BH; ++h) {
		const int x0 = 16 * cb + BH * h;
		__syncthreads();
#pragma unroll
		for (int x = 0; x < BH; ++x) {
			E[x * 64] = acc[x0 + x][r];
			if (NSET == 2)
				E[(BH + x) * 64] = bload_ae<AE>(wr, wvo,
				    (r * CW + x0 + x) * WST);
		}
		if (!act)
			continue;
		const int arow = FM::ROW_F64 ? (lg + 4 * r) : (4 * lg + r);
		const int a = a_lo + arow;
		const int c = c_lo + 16 * cb + ln;
		if (a < na && c < nc)
			epi_block<F, KIND, NSET, BH, STG ? TC : 0, SBA>(p, i, j, k, a,
			    c, b0u + BH * h, E, BH, sbc + 16 * cb + BH * h * TC,
			    sba + arow + BH * h * 32, sbv + BH * h, esum, ecmp);
	}

	/* deterministic workgroup reduction */
	double e = esum - ecmp;
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		e += __shfl_xor(e, off);
	if (lane == 0)
		red[wave] = e;
	__syncthreads();
	if (tid == 0) {
		double t = (red[0] + red[1]) + (red[2] + red[3]);
		if constexpr (NW == 8)
			t += (red[4] + red[5]) + (red[6] + red[7]);
		p.partial[it] = t;
	}
}

/* save area `slot` of wsave (two W sets; pt_wsave_bytes each) */
template <int F, int NW>
__device__ __forceinline__ typename Form<F>::AE *
wsave_of(void *wsave, int64_t slot)
{
	constexpr int64_t bytes = (int64_t)NW * Geo<F, 2, NW>::CW * 4 * 64 *
	    sizeof(typename Form<F>::AE);
	return (typename Form<F>::AE *)((char *)wsave + slot * bytes);
}

/*
 * One launch over items [item0, item0 + n_chunk) of one spin case (item =
 * ijk * n_tiles + tile).  XCD-aware order: workgroup n runs on XCD n % 8,
 * and each XCD gets a contiguous run of the launch's items, so the
 * workgroups running together on one XCD share operand slices in L2.  With
 * kgrp > 1 that run is taken in groups of kgrp consecutive ijk, tile-major
 * within a group: the work list holds (i, j, k) with k fastest, and
 * consecutive k share the slabs of i and j (six of the eight ovvv terms of
 * the mixed-spin case), which the kgrp items of one tile then read
 * together.  Each item still writes its own energy at its place in the
 * work list, so the energies do not depend on kgrp.  One workgroup per
 * item, handed out by the dispatcher in order; with two W sets item n of
 * the launch keeps W of set 0 in save area n (n_chunk <= p.nwsave).
 */
/* work-list item of launch position it (kgrp > 1: groups of kgrp ijk,
 * tile-major within a group) */
__device__ __forceinline__ int64_t
item_of(const pt_launch &p, int64_t it)
{
	if (p.kgrp <= 1)
		return p.item0 + it;
	const int64_t nt = p.n_tiles, G = p.kgrp, m = p.n_chunk / nt;
	const int64_t grp = it / (G * nt), r = it - grp * G * nt;
	const int64_t gsz = G < m - grp * G ? G : m - grp * G;
	const int64_t tile = r / gsz, g = r - tile * gsz;
	return p.item0 + (grp * G + g) * nt + tile;
}

template <int F, int KIND, int NSET, int NW>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 1 : 2) void
k_tiles(pt_launch p)
{
	const int64_t per = (p.n_chunk + 7) / 8;
	const int64_t x = blockIdx.x & 7;
	const int64_t it = x * per + (blockIdx.x >> 3);
	if (it >= p.n_chunk)
		return;
	tile_item<F, KIND, NSET, NW>(p, item_of(p, it),
	    NSET == 2 ? wsave_of<F, NW>(p.wsave, it) : nullptr);
}

/*
 * The same items from a persistent grid: workgroup w serves XCD w % 8 and
 * takes that XCD's run of launch positions, [x per, x per + per), one at a
 * time from the XCD's counter, in order.  The counter is fetched by one
 * thread between two barriers that every wave meets outside any divergent
 * branch (the previous item's partial-energy store is behind the first),
 * the value is read back as a wave-uniform scalar, and the loop ends for
 * the whole workgroup when it passes the run's end, which the host bounds;
 * no workgroup waits on another.  With two W sets each workgroup keeps W of
 * set 0 in its own save area (slot w).
 */
template <int F, int KIND, int NSET, int NW>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 1 : 2) void
k_tiles_p(pt_launch p)
{
	__shared__ int s_next;
	const int64_t per = (p.n_chunk + 7) / 8;
	const int x = blockIdx.x & 7;
	const int64_t lo = x * per;
	const int64_t cnt = p.n_chunk - lo < per ? p.n_chunk - lo : per;
	for (;;) {
		__syncthreads();
		if (threadIdx.x == 0)
			s_next = (int)atomicAdd(p.ctr + x, 1u);
		__syncthreads();
		const int64_t r = __builtin_amdgcn_readfirstlane(s_next);
		if (r >= cnt)
			break;
		tile_item<F, KIND, NSET, NW>(p, item_of(p, lo + r),
		    NSET == 2 ? wsave_of<F, NW>(p.wsave, blockIdx.x) : nullptr);
	}
}

/* the 8 item counters of a persistent launch, zeroed on its stream by a
 * kernel (a graph may record it) */
__global__ void
k_zero_ctr(uint32_t *c)
{
	if (threadIdx.x < 8)
		c[threadIdx.x] = 0u;
}

/* fixed-order compensated sum of x[0..n) by one workgroup of 1024 */
__device__ __forceinline__ void
reduce_block(const double *x, int64_t n, double *out)
{
	__shared__ double sh[1024];
	double s = 0.0, c = 0.0;
	for (int64_t t = threadIdx.x; t < n; t += 1024)
		kahan(s, c, x[t]);
	sh[threadIdx.x] = s - c;
	__syncthreads();
	for (int w = 512; w >= 1; w >>= 1) {
		if ((int)threadIdx.x < w)
			sh[threadIdx.x] += sh[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		out[0] = sh[0];
}

/* one workgroup, fixed-order compensated sum of x[0..n) */
__global__ __launch_bounds__(1024) void
k_reduce(const double *x, int64_t n, double *out)
{
	reduce_block(x, n, out);
}

/* one workgroup per segment: out[s] = the k_reduce sum of its partials */
__global__ __launch_bounds__(1024) void
k_reduce_batch(const pt_launch *segs, const int64_t *begin, double *out)
{
	const int s = blockIdx.x;
	reduce_block(segs[s].partial, begin[s + 1] - begin[s], out + s);
}

/* out[s] = sum of segment s (in order) */
__global__ void
k_segsum(const double *x, int64_t nseg, int64_t seglen, double *out)
{
	int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nseg)
		return;
	double acc = 0.0, c = 0.0;
	for (int64_t t = 0; t < seglen; ++t)
		kahan(acc, c, x[s * seglen + t]);
	out[s] = acc - c;
}

__device__ __forceinline__ void
split2(double x, int fp16, uint16_t &hb, uint16_t &lb)
{
	if (fp16) {
		_Float16 h = (_Float16)x;
		_Float16 l = (_Float16)(x - (double)h);
		hb = __builtin_bit_cast(uint16_t, h);
		lb = __builtin_bit_cast(uint16_t, l);
	} else {
		__bf16 h = (__bf16)(float)x;
		__bf16 l = (__bf16)(float)(x - (double)(float)h);
		hb = __builtin_bit_cast(uint16_t, h);
		lb = __builtin_bit_cast(uint16_t, l);
	}
}

/* [R][K][PL] -> [R][nkb][PL][KC] (FP64/FP32) or [R][nkb][PL][hi KC][lo KC]
 * (split, of x * scale); kappa >= K is zero.  P1 > 0: the plane PL = P1 x
 * P2 ([x][y], y contiguous) is stored transposed, [y][x] with x
 * contiguous.  One thread per group, each group written contiguously. */
__global__ void
k_pack(const void *src, void *dst, int64_t R, int64_t K, int64_t nkb,
    int64_t PL, int KC, int mode, double scale, int64_t P1)
{
	const int64_t total = R * nkb * PL;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	    t < total; t += (int64_t)gridDim.x * blockDim.x) {
		const int64_t pd = t % PL, rk = t / PL, kb = rk % nkb, r = rk / nkb;
		/* destination plane index pd -> source plane index pl */
		const int64_t pl = P1 > 0 ? (pd % P1) * (PL / P1) + pd / P1 : pd;
		const int64_t k0 = kb * KC;
		const int64_t sbase = (r * K + k0) * PL + pl;
		if (mode == PT_FORM_FP64) {
			const double *sp = (const double *)src + sbase;
			double *dp = (double *)dst + t * KC;
			for (int kk = 0; kk < KC; ++kk)
				dp[kk] = k0 + kk < K ? sp[(int64_t)kk * PL] : 0.0;
		} else if (mode == PT_FORM_FP32) {
			const float *sp = (const float *)src + sbase;
			float *dp = (float *)dst + t * KC;
			for (int kk = 0; kk < KC; ++kk)
				dp[kk] = k0 + kk < K ? sp[(int64_t)kk * PL] : 0.0f;
		} else {
			/* per kappa pair q: hi 2q, hi 2q + 1, lo 2q, lo 2q + 1 */
			const double *sp = (const double *)src + sbase;
			uint16_t *dp = (uint16_t *)dst + t * 2 * KC;
			for (int kk = 0; kk < KC; ++kk) {
				uint16_t hb = 0, lb = 0;
				if (k0 + kk < K)
					split2(sp[(int64_t)kk * PL] * scale,
					    mode == PT_FORM_F16X3, hb, lb);
				dp[4 * (kk >> 1) + (kk & 1)] = hb;
				dp[4 * (kk >> 1) + 2 + (kk & 1)] = lb;
			}
		}
	}
}

/* hi + lo split of (x * scale) into bf16 or fp16 pairs */
__global__ void
k_split(const double *src, uint32_t *dst, int64_t n, int fp16, double scale)
{
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x) {
		double x = src[t] * scale;
		uint16_t hb, lb;
		if (fp16) {
			_Float16 h = (_Float16)x;
			_Float16 l = (_Float16)(x - (double)h);
			hb = __builtin_bit_cast(uint16_t, h);
			lb = __builtin_bit_cast(uint16_t, l);
		} else {
			__bf16 h = (__bf16)(float)x;
			__bf16 l = (__bf16)(float)(x - (double)(float)h);
			hb = __builtin_bit_cast(uint16_t, h);
			lb = __builtin_bit_cast(uint16_t, l);
		}
		dst[t] = (uint32_t)hb | ((uint32_t)lb << 16);
	}
}

/* max |x| as the bit pattern of a non-negative double (order-free) */
__global__ void
k_absmax(const double *src, int64_t n, unsigned long long *out)
{
	__shared__ double sh[256];
	double m = 0.0;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x)
		m = fmax(m, fabs(src[t]));
	sh[threadIdx.x] = m;
	__syncthreads();
	for (int w = 128; w >= 1; w >>= 1) {
		if ((int)threadIdx.x < w)
			sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		atomicMax(out, (unsigned long long)__double_as_longlong(sh[0]));
}

/* plain copy (traffic counter calibration) */
__global__ void
k_copy(const double *src, double *dst, int64_t n)
{
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x)
		dst[t] = src[t];
}

__global__ void
k_synth(pt_synth_spec spec, int tensor, void *dst, int64_t n, int fp32)
{
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
	    t += (int64_t)gridDim.x * blockDim.x) {
		double x = pt_synth_elem(&spec, tensor, t);
		if (fp32)
			((float *)dst)[t] = (float)x;
		else
			((double *)dst)[t] = x;
	}
}

template <int F, int NSET, int NW>
int
launch_kind(const pt_launch *p, hipStream_t s)
{
	typedef Geo<F, NSET, NW> GM;
	if (p->tcw != GM::TC)
		return (int)hipErrorInvalidValue;
	dim3 block(GM::THREADS);
	if (p->ctr) {
		/* persistent grid: one save area per workgroup */
		if (p->pgrid <= 0 || p->pgrid % 8 || p->n_chunk > INT32_MAX / 2 ||
		    (NSET == 2 && (!p->wsave || p->pgrid > p->nwsave)))
			return (int)hipErrorInvalidValue;
		k_zero_ctr<<<1, 64, 0, s>>>(p->ctr);
		dim3 grid((unsigned)p->pgrid);
		if (p->kind == PT_AAA)
			k_tiles_p<F, PT_AAA, NSET, NW><<<grid, block, 0, s>>>(*p);
		else
			k_tiles_p<F, PT_AAB, NSET, NW><<<grid, block, 0, s>>>(*p);
		return (int)hipGetLastError();
	}
	/* one save area per item */
	if (NSET == 2 && (!p->wsave || p->n_chunk > p->nwsave))
		return (int)hipErrorInvalidValue;
	dim3 grid((unsigned)(8 * ((p->n_chunk + 7) / 8)));
	if (p->kind == PT_AAA)
		k_tiles<F, PT_AAA, NSET, NW><<<grid, block, 0, s>>>(*p);
	else
		k_tiles<F, PT_AAB, NSET, NW><<<grid, block, 0, s>>>(*p);
	return (int)hipGetLastError();
}

template <int F>
int
launch_form(const pt_launch *p, hipStream_t s)
{
	if (p->nw == 8)
		return p->nset == 1 ? launch_kind<F, 1, 8>(p, s) :
		    launch_kind<F, 2, 8>(p, s);
	return p->nset == 1 ? launch_kind<F, 1, 4>(p, s) :
	    launch_kind<F, 2, 4>(p, s);
}

} /* namespace */

extern "C" int
pt_launch_tiles(int form, const struct pt_launch *p, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	if (p->n_chunk <= 0)
		return 0;
	switch (form) {
	case PT_FORM_FP64: return launch_form<PT_FORM_FP64>(p, s);
	case PT_FORM_FP32: return launch_form<PT_FORM_FP32>(p, s);
	case PT_FORM_BF16X3: return launch_form<PT_FORM_BF16X3>(p, s);
	case PT_FORM_F16X3: return launch_form<PT_FORM_F16X3>(p, s);
	}
	return (int)hipErrorInvalidValue;
}

extern "C" int
pt_launch_reduce_batch(const struct pt_launch *segs, const int64_t *begin,
    int nseg, double *out, void *stream)
{
	if (nseg <= 0)
		return 0;
	k_reduce_batch<<<nseg, 1024, 0, (hipStream_t)stream>>>(segs, begin, out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_reduce(const double *x, int64_t n, double *out, void *stream)
{
	k_reduce<<<1, 1024, 0, (hipStream_t)stream>>>(x, n, out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_segsum(const double *x, int64_t nseg, int64_t seglen, double *out,
    void *stream)
{
	if (nseg <= 0)
		return 0;
	k_segsum<<<(unsigned)((nseg + 255) / 256), 256, 0, (hipStream_t)stream>>>(
	    x, nseg, seglen, out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_split(const double *src, uint32_t *dst, int64_t n, int fp16,
    double scale, void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_split<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(src, dst, n, fp16, scale);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_pack(const void *src, void *dst, int64_t R, int64_t K, int64_t nkb,
    int64_t PL, int KC, int mode, double scale, int64_t P1, void *stream)
{
	const int64_t total = R * nkb * PL;
	if (total <= 0)
		return 0;
	int64_t nb = (total + 255) / 256;
	k_pack<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(src, dst, R, K, nkb, PL, KC, mode, scale, P1);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_absmax(const double *src, int64_t n, double *out, void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_absmax<<<(unsigned)(nb < 4096 ? nb : 4096), 256, 0,
	    (hipStream_t)stream>>>(src, n, (unsigned long long *)out);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_copy(const double *src, double *dst, int64_t n, void *stream)
{
	if (n <= 0)
		return 0;
	k_copy<<<16384, 256, 0, (hipStream_t)stream>>>(src, dst, n);
	return (int)hipGetLastError();
}

extern "C" int
pt_launch_synth(const void *spec, int tensor, void *dst, int64_t n, int fp32,
    void *stream)
{
	if (n <= 0)
		return 0;
	int64_t nb = (n + 255) / 256;
	k_synth<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(*(const pt_synth_spec *)spec, tensor, dst, n,
	    fp32);
	return (int)hipGetLastError();
}
