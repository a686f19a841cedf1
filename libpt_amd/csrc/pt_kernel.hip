/*
 * CDNA4 (gfx950) kernels of the perturbative-triples engine.
 *
 * k_tiles: one workgroup = one ijk item x one (A, B, C) tile of 32 a x 32 b
 * x TCW c (TCW = 32 for (T), 16 for (fT)).  Four waves each own 16 x 16 x TCW
 * (a, b, c) block of the connected triples W in MFMA accumulators and never
 * write W to memory:
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
 *       one deterministic reduction per workgroup into partial[item].
 *
 * Operand tiles are staged global -> LDS per kappa chunk with the
 * contiguous global index mapped to consecutive threads.
 *
 * Forms: FP64 (v_mfma_f64_16x16x4_f64), FP32 (v_mfma_f32_16x16x4_f32),
 * split BF16 / FP16 (hi+lo pairs, hi*hi + hi*lo + lo*hi with
 * v_mfma_f32_16x16x16_{bf16,f16}, FP32 accumulation, FP64 epilogue).
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

#ifndef PT_KERNEL_AGPR
#define PT_KERNEL_AGPR 0
#endif

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));

namespace {

template <int F> struct Form;

template <> struct Form<PT_FORM_FP64> {
	typedef double G;   /* operand element in HBM */
	typedef double L;   /* LDS element (one plane) */
	typedef double AE;  /* accumulator element */
	typedef v4d ACC;
	typedef double ER;  /* epilogue real */
	static constexpr int NP = 1, KC = 8, KS = 4;
	static constexpr bool ROW_F64 = true, SPLIT = false;
};

template <> struct Form<PT_FORM_FP32> {
	typedef float G;
	typedef float L;
	typedef float AE;
	typedef v4f ACC;
	typedef float ER;
	static constexpr int NP = 1, KC = 16, KS = 4;
	static constexpr bool ROW_F64 = false, SPLIT = false;
};

template <> struct Form<PT_FORM_BF16X3> {
	typedef uint32_t G;  /* hi in bits 0-15, lo in bits 16-31 */
	typedef uint32_t L;  /* LDS word: two hi or two lo halves */
	typedef float AE;
	typedef v4f ACC;
	typedef double ER;
	static constexpr int NP = 1, KC = 16, KS = 16;
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

/* one K=16 bf16 / f16 MFMA on operands given as two 32-bit words each
 * (halves k0, k1 and k2, k3 of the lane's kappa) */
template <int F>
__device__ __forceinline__ v4f
mma16(uint32_t a01, uint32_t a23, uint32_t b01, uint32_t b23, v4f c)
{
	const v2u a = {a01, a23}, b = {b01, b23};
	if constexpr (F == PT_FORM_BF16X3)
		return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
		    __builtin_bit_cast(v4s, a), __builtin_bit_cast(v4s, b), c, 0, 0,
		    0);
	else
		return __builtin_amdgcn_mfma_f32_16x16x16f16(
		    __builtin_bit_cast(v4h, a), __builtin_bit_cast(v4h, b), c, 0, 0,
		    0);
}

/* hi*hi + hi*lo + lo*hi (FP32 accumulation) of one split operand pair.
 * Operand words (LDS layout of the split forms): w0 = hi k0 k1,
 * w1 = lo k0 k1, w2 = hi k2 k3, w3 = lo k2 k3; the small products first. */
template <int F>
__device__ __forceinline__ v4f
mma_split(v4u a, v4u b, v4f c)
{
	c = mma16<F>(a[1], a[3], b[0], b[2], c);
	c = mma16<F>(a[0], a[2], b[1], b[3], c);
	c = mma16<F>(a[0], a[2], b[0], b[2], c);
	return c;
}

template <class G>
__device__ __forceinline__ G
neg_if(G v, int neg)
{
	if constexpr (sizeof(G) == 4 && !__is_same(G, float))
		return neg ? (v ^ 0x80008000u) : v; /* both halves */
	else
		return neg ? -v : v;
}

/* ---------------------------------------------------------- epilogue */

template <class ER>
__device__ __forceinline__ ER
z_same(const ER *u, const ER *g, int64_t o, int64_t v, int i, int j, int k,
    int a, int b, int c)
{
	/* src/pt.c:162-175 */
	const int64_t vv = v * v;
	const ER *gjk = g + ((int64_t)j * o + k) * vv;
	const ER *gik = g + ((int64_t)i * o + k) * vv;
	const ER *gji = g + ((int64_t)j * o + i) * vv;
	return +u[i * v + a] * gjk[b * v + c]
	       -u[j * v + a] * gik[b * v + c]
	       -u[k * v + a] * gji[b * v + c]
	       -u[i * v + b] * gjk[a * v + c]
	       +u[j * v + b] * gik[a * v + c]
	       +u[k * v + b] * gji[a * v + c]
	       -u[i * v + c] * gjk[b * v + a]
	       +u[j * v + c] * gik[b * v + a]
	       +u[k * v + c] * gji[b * v + a];
}

template <class ER>
__device__ __forceinline__ ER
z_mixed(const pt_epi &e, int64_t oa, int64_t va, int64_t ob, int64_t vb,
    int i, int j, int k, int a, int b, int c)
{
	/* src/pt.c:177-185 and :433-442 */
	const ER *u1 = (const ER *)e.u1, *u2 = (const ER *)e.u2;
	const ER *f1 = (const ER *)e.f1, *f2 = (const ER *)e.f2;
	const ER *GA = (const ER *)e.GA, *GB = (const ER *)e.GB;
	const ER *ZA = (const ER *)e.ZA, *ZB = (const ER *)e.ZB;
	const int64_t jk = ((int64_t)j * ob + k) * va, ik = ((int64_t)i * ob + k) * va;
	const int64_t jkbc = (jk + b) * vb + c, jkac = (jk + a) * vb + c;
	const int64_t ikbc = (ik + b) * vb + c, ikac = (ik + a) * vb + c;
	const int64_t jiba = (((int64_t)j * oa + i) * va + b) * va + a;
	return (u1[i * va + a] * GB[jkbc] + f1[i * va + a] * ZB[jkbc])
	     - (u1[i * va + b] * GB[jkac] + f1[i * va + b] * ZB[jkac])
	     - (u1[j * va + a] * GB[ikbc] + f1[j * va + a] * ZB[ikbc])
	     + (u1[j * va + b] * GB[ikac] + f1[j * va + b] * ZB[ikac])
	     + (u2[k * vb + c] * GA[jiba] + f2[k * vb + c] * ZA[jiba]);
}

__device__ __forceinline__ void
kahan(double &s, double &cmp, double x)
{
	double y = x - cmp;
	double t = s + y;
	cmp = (t - s) - y;
	s = t;
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
	int role, set, K, neg, packed, ufast;
	int s0, sn, sext, u0, uext, w0, wext, qwe, qsh;
	int lo_mask, hi_sh;  /* pair index split: fast index bits, shift */
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
	s.neg = uni(T.sign) < 0;
	s.packed = uni(T.packed);
	s.ufast = uni(T.ufast);
	s.s_st = uni64(T.s_st);
	s.qgs = uni64(T.q_k) * GB;
	const int na = p.na, nc = p.nc;
	int ur, wr;
	if (s.role == PT_ROLE_A) {
		s.s0 = A0; s.sn = 32; s.sext = na;
		s.u0 = B0; s.uext = na; s.w0 = C0; s.wext = nc; s.qwe = tcw;
		ur = PT_ROLE_B; wr = PT_ROLE_C;
	} else if (s.role == PT_ROLE_B) {
		s.s0 = B0; s.sn = 32; s.sext = na;
		s.u0 = A0; s.uext = na; s.w0 = C0; s.wext = nc; s.qwe = tcw;
		ur = PT_ROLE_A; wr = PT_ROLE_C;
	} else {
		s.s0 = C0; s.sn = tcw; s.sext = nc;
		s.u0 = A0; s.uext = na; s.w0 = B0; s.wext = na; s.qwe = 32;
		ur = PT_ROLE_A; wr = PT_ROLE_B;
	}
	s.qsh = s.qwe == 32 ? 5 : 4;
	s.lo_mask = s.ufast ? 31 : s.qwe - 1;
	s.hi_sh = s.ufast ? 5 : s.qsh;
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

/* (u, w) of pair index pidx: the contiguous index of Q varies fastest */
__device__ __forceinline__ void
pair_uw(const Step &s, int pidx, int &u, int &w)
{
	const int lo = pidx & s.lo_mask, hi = pidx >> s.hi_sh;
	u = s.ufast ? lo : hi;
	w = s.ufast ? hi : lo;
}

/* raw buffer loads: offsets at or beyond the descriptor's byte count
 * return zero, which handles tile edges, invalid pairs and the kappa tail
 * without branches */
template <class G>
__device__ __forceinline__ G
bload(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	if constexpr (sizeof(G) == 8)
		return __builtin_bit_cast(G,
		    __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
	else
		return __builtin_bit_cast(G,
		    __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

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
 * FP64 / FP32 staging: NSLOT slots, each holding one step (term, kappa
 * chunk) of the tile, loaded NSLOT - 1 steps ahead of the step computed.
 * Slot (bytes):
 *   S: 32 rows of 64 B of data (the chunk's kappa of one row) + padding;
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
 * lane n of a lane group reads row n.  The row strides are those at which
 * such reads have no LDS bank conflicts on gfx950 (measured with
 * SQ_LDS_BANK_CONFLICT): 16-byte reads over 16 rows need a stride of 32
 * modulo 64 bytes (type A: S 96, Q 1056), 8-byte reads 16 modulo 64 (type
 * B: S 80, Q 1040); the other strides cost a 2-way conflict.
 */
struct Stg {
	static constexpr int S_ROW_A = 96, S_ROW_B = 80;
	static constexpr int Q_USTR_A = 1024 + 32, Q_USTR_B = 1024 + 16;
	static constexpr int S_BYTES = 32 * S_ROW_A;
	static constexpr int Q_BYTES = 32 * Q_USTR_A;
	static constexpr int SLOT = S_BYTES + Q_BYTES;
	static constexpr int NSLOT = 2;
	static constexpr int BYTES = NSLOT * SLOT;
	template <bool TA> static constexpr int s_row() {
		return TA ? S_ROW_A : S_ROW_B;
	}
	template <bool TA> static constexpr int q_ustr() {
		return TA ? Q_USTR_A : Q_USTR_B;
	}
};

#define PT_QRUN_W 8   /* DMA runs per wave per step (32 u-rows / 4 waves) */

/* lane offsets of the Q runs of one term (type A: w = lane / 4, 16-byte
 * chunk c = lane % 4 from group c / 2; type B: w = lane / 2, chunk
 * lane % 2); gstride = bytes per packed group of the slab */
template <bool TA>
__device__ __forceinline__ void
qtab(const Step &s, int wave, int lane, uint32_t (&vo)[PT_QRUN_W])
{
	const int w = TA ? (lane >> 2) : (lane >> 1);
	const uint32_t cofs = TA ? (uint32_t)((lane & 1) * 16 +
	    ((lane >> 1) & 1) * s.qgs) : (uint32_t)((lane & 1) * 16);
#pragma unroll
	for (int n = 0; n < PT_QRUN_W; ++n) {
		const int ug = s.u0 + wave + 4 * n, wg = s.w0 + w;
		const bool v = ug < s.uext && wg < s.wext && (!s.packed || ug > wg);
		const int64_t pidx = s.packed ? (int64_t)ug * (ug - 1) / 2 + wg :
		    (int64_t)ug * s.q_u + (int64_t)wg * s.q_w;
		vo[n] = v ? (uint32_t)(pidx * 32) + cofs : PT_OOR;
	}
}

/* DMA of step g (kappa groups g or 2g, 2g + 1) of a term into a slot */
template <bool TA>
__device__ __forceinline__ void
issue_q(const Step &s, int g, char *sQ, int wave,
    const uint32_t (&vo)[PT_QRUN_W])
{
	constexpr int NG = TA ? 2 : 1;
	const __amdgpu_buffer_rsrc_t rq = brsrc(s.qb0 + (int64_t)g * NG * s.qgs,
	    NG * s.qgs);
#pragma unroll
	for (int n = 0; n < PT_QRUN_W; ++n)
		__builtin_amdgcn_raw_ptr_buffer_load_lds(rq,
		    (__attribute__((address_space(3))) void *)(sQ +
		    (wave + 4 * n) * Stg::q_ustr<TA>()), 16, vo[n], 0, 0, 0);
}

/* wait until at most n vector-memory loads of this wave are outstanding */
template <int N>
__device__ __forceinline__ void
wait_vm()
{
	__asm__ volatile("s_waitcnt vmcnt(%0)" :: "n"(N) : "memory");
}

/* S loads: thread -> (row, kappa k0 .. k0 + SEPT - 1) of a chunk of KS
 * kappa, KS / SEPT threads per row (8 for type A, 4 for type B); rows
 * beyond the tile load zeros and are not stored.  Per term: the byte
 * offset from the chunk base (or PT_OOR) and the LDS store offset (or -1). */
template <int F, int SEPT>
__device__ __forceinline__ void
s_lane(const Step &s, int tid, bool ta, uint32_t &ofs, int &lofs)
{
	constexpr int ES = (int)sizeof(typename Form<F>::G);
	const int tpr = ta ? 8 : 4;
	const int row = tid / tpr, k0 = (tid % tpr) * SEPT;
	const int sg = s.s0 + row;
	ofs = (row < s.sn && sg < s.sext) ?
	    (uint32_t)(((int64_t)sg * s.s_st + k0) * ES) : PT_OOR;
	lofs = row < 32 ? row * (ta ? Stg::S_ROW_A : Stg::S_ROW_B) + k0 * ES :
	    -1;
}

template <int F, int SEPT>
__device__ __forceinline__ void
load_s(const Step &s, uint32_t ofs, int k0, int kc, int ks,
    typename Form<F>::G (&sr)[SEPT])
{
	typedef typename Form<F>::G G;
	constexpr int ES = (int)sizeof(G);
	const __amdgpu_buffer_rsrc_t rs = brsrc(s.sb0 + (int64_t)kc * ES,
	    s.sbytes - (int64_t)kc * ES);
#pragma unroll
	for (int e = 0; e < SEPT; ++e) {
		const uint32_t o = (k0 + e < ks && kc + k0 + e < s.K &&
		    ofs != PT_OOR) ? ofs + e * ES : PT_OOR;
		sr[e] = neg_if(bload<G>(rs, o), s.neg);
	}
}

template <int F, int SEPT>
__device__ __forceinline__ void
store_s(char *sS, int lofs, const typename Form<F>::G (&sr)[SEPT])
{
	typedef typename Form<F>::G G;
	if (lofs < 0)
		return;
	if constexpr (Form<F>::SPLIT) {
		/* two kappa (hi | lo << 16 each) -> hi pair word, lo pair word */
		static_assert(SEPT == 2, "split S pairs");
		const uint32_t h = (sr[0] & 0xffffu) | (sr[1] << 16);
		const uint32_t l = (sr[0] >> 16) | (sr[1] & 0xffff0000u);
		*(v2u *)(sS + lofs) = v2u{h, l};
	} else {
#pragma unroll
		for (int e = 0; e < SEPT; ++e)
			*(G *)(sS + lofs + e * (int)sizeof(G)) = sr[e];
	}
}

/* a split-form operand from LDS: 16 bytes (type A: hi k0 k1, lo k0 k1,
 * hi k2 k3, lo k2 k3) or 8 bytes (type B: hi k0 k1, lo k0 k1, with k2 and
 * k3 zero in both operands) */
template <bool TA>
__device__ __forceinline__ v4u
split_rd(const char *p)
{
	if constexpr (TA)
		return *(const v4u *)p;
	else {
		const v2u t = *(const v2u *)p;
		return v4u{t[0], t[1], 0u, 0u};
	}
}

/* an MFMA operand chunk of NB bytes (16: type A, 8: type B) */
template <class L, int NB>
struct Chunk {
	typedef L V __attribute__((ext_vector_type(NB / sizeof(L))));
	static constexpr int N = NB / sizeof(L);
};

/*
 * Phase 1 (accumulator block x = c offset, D[a][b]) over one staged step:
 *   role a (RA): A = S[a] (fixed), B = Q[b][c = x];
 *   role b:      A = Q[a][c = x],  B = S[b] (fixed).
 * TA: type-A rows (64 B per w), else type-B rows (32 B per w).
 */
template <int F, int TCW, bool RA, bool TA>
__device__ __forceinline__ void
mma_p1(typename Form<F>::ACC (&ac)[TCW], const char *sS, const char *sQ,
    int wa, int wb, int ln, int lg)
{
	typedef typename Form<F>::L L;
	constexpr int NB = TA ? 16 : 8, WB = TA ? 64 : 32;
	const int ws = RA ? wa : wb, wu = RA ? wb : wa;
	const char *pf = sS + (16 * ws + ln) * Stg::s_row<TA>() + NB * lg;
	const char *pv = sQ + (16 * wu + ln) * Stg::q_ustr<TA>() + NB * lg;
	if constexpr (Form<F>::SPLIT) {
		const v4u f = split_rd<TA>(pf);
#pragma unroll
		for (int x = 0; x < TCW; ++x) {
			const v4u q = split_rd<TA>(pv + WB * x);
			ac[x] = RA ? mma_split<F>(f, q, ac[x]) :
			    mma_split<F>(q, f, ac[x]);
		}
	} else {
		typedef typename Chunk<L, NB>::V V;
		constexpr int NE = Chunk<L, NB>::N;
		const V f = *(const V *)pf;
#pragma unroll
		for (int x = 0; x < TCW; ++x) {
			const V q = *(const V *)(pv + WB * x);
#pragma unroll
			for (int e = 0; e < NE; ++e)
				ac[x] = RA ? mma(f[e], q[e], ac[x]) :
				    mma(q[e], f[e], ac[x]);
		}
	}
}

/*
 * Phase 2 (block x = 16 cb + b, D[a][c]): role c: A = Q[a][b], B = S[c];
 * always type-B rows.
 */
template <int F, int TCW>
__device__ __forceinline__ void
mma_p2(typename Form<F>::ACC (&ac)[TCW], const char *sS, const char *sQ,
    int wa, int wb, int ln, int lg)
{
	typedef typename Form<F>::L L;
	constexpr int NCB = TCW / 16;
	const char *pv = sQ + (16 * wa + ln) * Stg::Q_USTR_B + 16 * wb * 32 +
	    8 * lg;
	if constexpr (Form<F>::SPLIT) {
		v4u f[NCB];
#pragma unroll
		for (int cb = 0; cb < NCB; ++cb)
			f[cb] = split_rd<false>(sS + (16 * cb + ln) * Stg::S_ROW_B +
			    8 * lg);
#pragma unroll
		for (int b = 0; b < 16; ++b) {
			const v4u q = split_rd<false>(pv + 32 * b);
#pragma unroll
			for (int cb = 0; cb < NCB; ++cb)
				ac[16 * cb + b] = mma_split<F>(q, f[cb], ac[16 * cb + b]);
		}
	} else {
		typedef typename Chunk<L, 8>::V V;
		constexpr int NE = Chunk<L, 8>::N;
		V f[NCB];
#pragma unroll
		for (int cb = 0; cb < NCB; ++cb)
			f[cb] = *(const V *)(sS + (16 * cb + ln) * Stg::S_ROW_B +
			    8 * lg);
#pragma unroll
		for (int b = 0; b < 16; ++b) {
			const V q = *(const V *)(pv + 32 * b);
#pragma unroll
			for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
				for (int e = 0; e < NE; ++e)
					ac[16 * cb + b] = mma(q[e], f[cb][e],
					    ac[16 * cb + b]);
		}
	}
}

/* -------------------------------------------------------------- kernel */

/* c tile width of a form and W-set count (as pt_tcw) */
constexpr int
tcw_of(int F, int NSET)
{
	return NSET > 0 && F == PT_FORM_FP32 ? 32 : 16;
}

/* Every tile kernel targets two workgroups per CU (PT_WG_PER_CU): one
 * workgroup's LDS (two slots, 76 KB) and registers (<= 256 per lane) allow
 * it, and one workgroup's staging and barriers then overlap the other's
 * MFMAs.  Two W sets keep one set of accumulators: W of set 0 waits in
 * the workgroup's save area while set 1 is formed. */

/*
 * One work item (ijk, tile) of a spin case: everything between reading the
 * item and writing its energy to partial[it].  KIND is the spin case,
 * PT_AAA or PT_AAB (only the epilogue depends on it).  With two W sets the
 * sets are formed one after the other; wsv is the item's save area for W
 * of set 0 (4 TCW values per lane per wave).
 */
template <int F, int KIND, int NSET>
__device__ __forceinline__ void
tile_item(const pt_launch &p, int64_t it, typename Form<F>::AE *wsv)
{
	typedef Form<F> FM;
	typedef typename FM::G G;
	typedef typename FM::L L;
	typedef typename FM::AE AE;
	typedef typename FM::ACC ACC;
	typedef typename FM::ER ER;
	constexpr int KC = FM::KC;
	constexpr int TCW = tcw_of(F, NSET);
	constexpr int NCB = TCW / 16;
	constexpr int GB = 32;  /* bytes per packed kappa group and pair */
	/* S elements per thread per step */
	constexpr int SEPT = F == PT_FORM_FP64 ? 1 : 2;
	constexpr int SMEM = Stg::BYTES;
	static_assert(SMEM >= 4 * NSET * TCW * 64 * (int)sizeof(AE),
	    "epilogue slots exceed the stage area");
	static_assert(NSET == 1 || NSET == 2, "one or two W sets");

	__shared__ __attribute__((aligned(16))) char smem[SMEM];
	__shared__ double red[4];
	__shared__ pt_term sterm[PT_MAX_TERMS];
	AE *sX = (AE *)smem;

	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wa = wave & 1, wb = wave >> 1;
	const int ln = lane & 15, lg = lane >> 4;

	const bool same = KIND == PT_AAA;
	const int64_t ijk_id = it / p.n_tiles;
	const int tile = (int)(it - ijk_id * p.n_tiles);
	const int i = p.ijk[3 * ijk_id], j = p.ijk[3 * ijk_id + 1];
	const int k = p.ijk[3 * ijk_id + 2];
	const int A0 = p.tiles[3 * tile] * 32, B0 = p.tiles[3 * tile + 1] * 32;
	const int C0 = p.tiles[3 * tile + 2] * TCW;
	const int na = p.na, nc = p.nc;
	const int a_lo = A0 + 16 * wa, b_lo = B0 + 16 * wb;
	/* does this wave's block hold any element of the triangle? */
	const bool act = a_lo < na && b_lo < na && a_lo + 15 > b_lo &&
	    (!same || b_lo + 15 > C0);

	ACC acc[TCW];
#pragma unroll
	for (int x = 0; x < TCW; ++x)
		acc[x] = ACC{0, 0, 0, 0};
	/* save area of W set 0: value (r, x) of (wave, lane) at element
	 * ((r TCW + x) 4 + wave) 64 + lane */
	constexpr int WST = 4 * 64 * (int)sizeof(AE);
	const __amdgpu_buffer_rsrc_t wr = brsrc(wsv, NSET == 2 ?
	    (int64_t)4 * TCW * WST : 0);
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
		for (int w = tid; w < nw; w += PT_THREADS)
			dst[w] = src[w];
		__syncthreads();
	}
	constexpr int ES = (int)sizeof(G);
	auto step_of = [&](int tn) {
		return make_step<ES, GB>(p, sterm[tn], i, j, k, A0, B0, C0, TCW);
	};

	/* [a][b]@c -> [a][c]@b through LDS scratch, per register r */
	auto transpose = [&](AE *scratch) {
#pragma unroll
		for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			AE *X = scratch + wave * (4 * 16 * 17);
			__syncthreads();
#pragma unroll
			for (int c = 0; c < 16; ++c)
				X[(lg * 16 + ln) * 17 + c] = acc[16 * cb + c][r];
			__syncthreads();
#pragma unroll
			for (int b = 0; b < 16; ++b)
				acc[16 * cb + b][r] = X[(lg * 16 + b) * 17 + ln];
		}
		__syncthreads();
	};

	{
		/*
		 * Steps n = 0 .. N-1 walk the terms (role a, role b, role c) and
		 * their kappa chunks; step n lives in slot n % NSLOT.  The loads of
		 * step n + D (D = NSLOT - 1: Q by DMA, S into registers) are issued
		 * while step n computes; S values wait in a register ring and are
		 * written to their slot one step before use.  One barrier per step.
		 */
		constexpr int NS = Stg::NSLOT, D = NS - 1;
		constexpr int KSA = KC, KSB = KC / 2;   /* kappa per step, type A / B */
		constexpr int VMS = PT_QRUN_W + SEPT;   /* loads per step per wave */
		const int wv = __builtin_amdgcn_readfirstlane(wave);
		const int nta = uni(p.nterm_a), nt1 = uni(p.nterm1);
		/* type A: a 16-wide w range (roles a and b of a 16-wide c tile) */
		auto type_a = [&](int role) { return TCW == 16 && role != PT_ROLE_C; };

		/* prefetch walker: term pt, chunk pk of step pn (the next to issue) */
		int pt = 0, pk = 0, pn = 0;
		Step pre = step_of(0);
		bool pa = type_a(pre.role);
		uint32_t vo[PT_QRUN_W], sofs;
		int slofs;
		auto term_tables = [&]() {
			if (pa)
				qtab<true>(pre, wv, lane, vo);
			else
				qtab<false>(pre, wv, lane, vo);
			s_lane<F, SEPT>(pre, tid, pa, sofs, slofs);
		};
		term_tables();
		/* register ring: R[m] = S of step n + 1 + m (m < D - 1), its LDS
		 * offset LR[m]; the loads of step n + D land in RN / LN */
		constexpr int DR = D > 1 ? D - 1 : 1;
		G R[DR][SEPT], RN[SEPT];
		int LR[DR], LN = -1;
#pragma unroll
		for (int m = 0; m < DR; ++m)
			LR[m] = -1;
		/* number of steps */
		int N = 0;
		for (int tt = 0; tt < nterm; ++tt) {
			const int kk = uni(sterm[tt].K);
			const int ks = type_a(uni(sterm[tt].role)) ? KSA : KSB;
			N += (kk + ks - 1) / ks;
		}
		/* issue step pn into its slot; its S goes to RN */
		auto issue = [&]() {
			char *slot = smem + (pn % NS) * Stg::SLOT;
			const int ks = pa ? KSA : KSB;
			if (pa)
				issue_q<true>(pre, pk / KSA, slot + Stg::S_BYTES, wv, vo);
			else
				issue_q<false>(pre, pk / KSB, slot + Stg::S_BYTES, wv, vo);
			const int k0 = (tid % (pa ? 8 : 4)) * SEPT;
			load_s<F, SEPT>(pre, sofs, k0, pk, ks, RN);
			LN = slofs;
			++pn;
			pk += ks;
			if (pk >= pre.K) {
				pk = 0;
				if (++pt < nterm) {
					pre = step_of(pt);
					pa = type_a(pre.role);
					term_tables();
				}
			}
		};
		/* after step n: the ring moves one step on, RN joins at the end */
		auto rotate = [&]() {
			if constexpr (D > 1) {
#pragma unroll
				for (int m = 0; m + 1 < DR; ++m) {
#pragma unroll
					for (int e = 0; e < SEPT; ++e)
						R[m][e] = R[m + 1][e];
					LR[m] = LR[m + 1];
				}
#pragma unroll
				for (int e = 0; e < SEPT; ++e)
					R[DR - 1][e] = RN[e];
				LR[DR - 1] = LN;
			}
		};
		/* wait until the loads of step m (issued) are complete: at most
		 * (pn - 1 - m) later steps' loads may stay outstanding */
		auto wait_step = [&](int m) {
			const int later = pn - 1 - m;
			if (D >= 3 && later >= 2)
				wait_vm<2 * VMS>();
			else if (D >= 2 && later >= 1)
				wait_vm<VMS>();
			else
				wait_vm<0>();
		};

		/* prologue: steps 0 .. D-1 in flight, step 0 staged; entering
		 * iteration n the ring holds R[m] = S of step n + 1 + m */
		{
			G r0[SEPT];
			issue();
#pragma unroll
			for (int e = 0; e < SEPT; ++e)
				r0[e] = RN[e];
			const int l0 = LN;
#pragma unroll
			for (int m = 0; m + 1 < D; ++m) {
				if (pn < N) {
					issue();
#pragma unroll
					for (int e = 0; e < SEPT; ++e)
						R[m][e] = RN[e];
					LR[m] = LN;
				}
			}
			wait_step(0);
			store_s<F, SEPT>(smem, l0, r0);
			__syncthreads();
		}

		/* compute walker: term ct, chunk ck of step n */
		int ct = 0, ck = 0, cK = uni(sterm[0].K);
		int n = 0;
		/* steps of phase PH up to term tend (exclusive) */
		auto run_phase = [&](auto ph, int tend) {
			constexpr int PH = decltype(ph)::value;
			constexpr bool TA = PH < 2 && TCW == 16;
			while (ct < tend) {
				/* loads of step n + D into the slot step n - 1 used */
				if (pn < N) {
					issue();
				}
				const char *slot = smem + (n % NS) * Stg::SLOT;
				const char *cS = slot, *cQ = slot + Stg::S_BYTES;
				if constexpr (PH < 2)
					mma_p1<F, TCW, PH == 0, TA>(acc, cS, cQ, wa, wb, ln,
					    lg);
				else
					mma_p2<F, TCW>(acc, cS, cQ, wa, wb, ln, lg);
				/* stage the S rows of step n + 1 */
				if (n + 1 < N) {
					wait_step(n + 1);
					char *sn = smem + ((n + 1) % NS) * Stg::SLOT;
					if constexpr (D > 1)
						store_s<F, SEPT>(sn, LR[0], R[0]);
					else
						store_s<F, SEPT>(sn, LN, RN);
					rotate();
				}
				__syncthreads();
				++n;
				ck += TA ? KSA : KSB;
				if (ck >= cK) {
					ck = 0;
					if (++ct < nterm)
						cK = uni(sterm[ct].K);
				}
			}
		};
#pragma unroll 1
		for (int s = 0; s < NSET; ++s) {
			const int t0 = s * nset_terms;
			run_phase(std::integral_constant<int, 0>(), t0 + nta);
			run_phase(std::integral_constant<int, 1>(), t0 + nt1);
			/* the slot of the last phase-1 step is free until step
			 * n + D issues into it: transpose through it */
			transpose((AE *)(smem + ((n + NS - 1) % NS) * Stg::SLOT));
			run_phase(std::integral_constant<int, 2>(), t0 + nset_terms);
			if (NSET == 2 && s == 0) {
				/* W of set 0 to the save area, lane-contiguous per
				 * (r, x, wave); the accumulators restart for set 1 */
#pragma unroll
				for (int r = 0; r < 4; ++r)
#pragma unroll
					for (int x = 0; x < TCW; ++x) {
						bstore_ae<AE>(wr, wvo, (r * TCW + x) * WST,
						    acc[x][r]);
						acc[x][r] = 0;
					}
			}
		}
	}

	/* epilogue: layout [a][c]@b, block x = 16 cb + b; lane holds
	 * c = C0 + 16 cb + ln.  Each register column r goes through this
	 * lane's own LDS slots so the element loop runs with a runtime index. */
	double esum = 0.0, ecmp = 0.0;
	AE *E = sX + wave * (NSET * TCW * 64) + lane;
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		__syncthreads();
#pragma unroll
		for (int x = 0; x < TCW; ++x) {
			E[x * 64] = acc[x][r];
			if (NSET == 2)
				E[(TCW + x) * 64] = bload_ae<AE>(wr, wvo,
				    (r * TCW + x) * WST);
		}
		if (!act)
			continue;
		const int arow = FM::ROW_F64 ? (lg + 4 * r) : (4 * lg + r);
		const int a = a_lo + arow;
		if (a >= na)
			continue;
		for (int x = 0; x < TCW; ++x) {
			const int b = b_lo + (x & 15);
			const int c = C0 + 16 * (x >> 4) + ln;
			if (!(c < nc && a > b && (!same || b > c)))
				continue;
			ER z, dn;
			const ER *d1 = (const ER *)p.epi.d1;
			if (same) {
				const int64_t o = p.oa, v = p.va;
				z = z_same<ER>((const ER *)p.epi.u1, (const ER *)p.epi.GA,
				    o, v, i, j, k, a, b, c) +
				    z_same<ER>((const ER *)p.epi.f1, (const ER *)p.epi.ZA,
				    o, v, i, j, k, a, b, c);
				dn = d1[i * v + a] + d1[j * v + b] + d1[k * v + c];
			} else {
				const ER *d2 = (const ER *)p.epi.d2;
				z = z_mixed<ER>(p.epi, p.oa, p.va, p.ob, p.vb, i, j, k, a, b,
				    c);
				dn = d1[i * p.va + a] + d1[j * p.va + b] + d2[k * p.vb + c];
			}
			ER term;
			if (NSET == 1) {
				ER w = (ER)E[x * 64];
				if (FM::SPLIT)
					w *= (ER)p.inv_scale[0];
				term = w * (w - z) / dn;
			} else {
				ER wl = (ER)E[(TCW + x) * 64], wr = (ER)E[x * 64];
				if (FM::SPLIT) {
					wl *= (ER)p.inv_scale[0];
					wr *= (ER)p.inv_scale[1];
				}
				term = (wl - z) * wr / dn;
			}
			kahan(esum, ecmp, (double)term);
		}
	}

	/* deterministic workgroup reduction */
	double e = esum - ecmp;
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		e += __shfl_xor(e, off);
	if (lane == 0)
		red[wave] = e;
	__syncthreads();
	if (tid == 0)
		p.partial[it] = (red[0] + red[1]) + (red[2] + red[3]);
}

/* save area `slot` of wsave (two W sets; pt_wsave_bytes each) */
template <int F>
__device__ __forceinline__ typename Form<F>::AE *
wsave_of(void *wsave, int64_t slot)
{
	constexpr int64_t bytes = (int64_t)4 * tcw_of(F, 2) * 4 * 64 *
	    sizeof(typename Form<F>::AE);
	return (typename Form<F>::AE *)((char *)wsave + slot * bytes);
}

/*
 * One launch over items [item0, item0 + n_chunk) of one spin case.  XCD-
 * aware order: workgroup n runs on XCD n % 8, and each XCD gets a
 * contiguous run of items, so neighbouring tiles of one ijk share L2.
 * One workgroup per item, handed out by the dispatcher in order; with two
 * W sets item item0 + n keeps W of set 0 in save area n (n_chunk <=
 * p.nwsave).
 */
template <int F, int KIND, int NSET>
__global__ __launch_bounds__(PT_THREADS, PT_WG_PER_CU) void
k_tiles(pt_launch p)
{
	const int64_t per = (p.n_chunk + 7) / 8;
	const int64_t x = blockIdx.x & 7;
	const int64_t it = x * per + (blockIdx.x >> 3);
	if (it < p.n_chunk)
		tile_item<F, KIND, NSET>(p, p.item0 + it,
		    NSET == 2 ? wsave_of<F>(p.wsave, it) : nullptr);
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
 * (split, of x * scale); kappa >= K is zero.  One thread per group: reads
 * coalesce over PL, each group is written contiguously. */
__global__ void
k_pack(const void *src, void *dst, int64_t R, int64_t K, int64_t nkb,
    int64_t PL, int KC, int mode, double scale)
{
	const int64_t total = R * nkb * PL;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	    t < total; t += (int64_t)gridDim.x * blockDim.x) {
		const int64_t pl = t % PL, rk = t / PL, kb = rk % nkb, r = rk / nkb;
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

template <int F, int NSET>
int
launch_kind(const pt_launch *p, hipStream_t s)
{
	/* one save area per item */
	if (NSET == 2 && (!p->wsave || p->n_chunk > p->nwsave))
		return (int)hipErrorInvalidValue;
	dim3 grid((unsigned)(8 * ((p->n_chunk + 7) / 8))), block(PT_THREADS);
	if (p->kind == PT_AAA)
		k_tiles<F, PT_AAA, NSET><<<grid, block, 0, s>>>(*p);
	else
		k_tiles<F, PT_AAB, NSET><<<grid, block, 0, s>>>(*p);
	return (int)hipGetLastError();
}

/* Kernels built in pt_kernel_agpr.hip, the same source compiled without
 * the VGPR-form MFMA option: FP64 with two W sets (the option gave wrong
 * energies when that kernel held both sets in registers; with the sets
 * formed one after the other it is correct there but no faster) and the
 * split forms (whose accumulators stay in place either way, and for which
 * the option can crash the compiler). */
constexpr bool
agpr_unit(int F, int NSET)
{
	return (F == PT_FORM_FP64 && NSET == 2) || F >= PT_FORM_BF16X3;
}

template <int F>
int
launch_form(const pt_launch *p, hipStream_t s)
{
	if (p->nset == 1) {
		if constexpr (!agpr_unit(F, 1) || PT_KERNEL_AGPR)
			return launch_kind<F, 1>(p, s);
	} else {
		if constexpr (!agpr_unit(F, 2) || PT_KERNEL_AGPR)
			return launch_kind<F, 2>(p, s);
	}
	return (int)hipErrorInvalidValue;
}

} /* namespace */

#if PT_KERNEL_AGPR
extern "C" int
pt_launch_tiles_agpr(int form, const struct pt_launch *p, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	switch (form) {
	case PT_FORM_FP64: return launch_kind<PT_FORM_FP64, 2>(p, s);
	case PT_FORM_BF16X3: return launch_form<PT_FORM_BF16X3>(p, s);
	case PT_FORM_F16X3: return launch_form<PT_FORM_F16X3>(p, s);
	}
	return (int)hipErrorInvalidValue;
}
#else
extern "C" int
pt_launch_tiles(int form, const struct pt_launch *p, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	if (p->n_chunk <= 0)
		return 0;
	if (agpr_unit(form, p->nset))
		return pt_launch_tiles_agpr(form, p, stream);
	switch (form) {
	case PT_FORM_FP64: return launch_form<PT_FORM_FP64>(p, s);
	case PT_FORM_FP32: return launch_form<PT_FORM_FP32>(p, s);
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
    int64_t PL, int KC, int mode, double scale, void *stream)
{
	const int64_t total = R * nkb * PL;
	if (total <= 0)
		return 0;
	int64_t nb = (total + 255) / 256;
	k_pack<<<(unsigned)(nb < 65536 ? nb : 65536), 256, 0,
	    (hipStream_t)stream>>>(src, dst, R, K, nkb, PL, KC, mode, scale);
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
#endif /* PT_KERNEL_AGPR */
