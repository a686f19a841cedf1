/*
 * Host side of the MI355X perturbative-triples engine: the pt.h drop-in
 * entry points (reference src/pt.h, src/pt.c:454-535, 936-1050,
 * src/ptmp.c same lines, src/misc.c) and the extension API of
 * libpt_amd.h.
 *
 * One call = one plan: stage inputs in HBM, build the work list of every
 * spin case (ijk items x (A, B) tiles, ijk-major), launch the tile kernel
 * over it, reduce the per-item energies in a fixed order, and combine the
 * spin cases with the reference's factors.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/libpt_amd.h"
#include "../../include/pt.h"
#include "pt_engine.h"
#include "pt_terms.h"
#include "synth.h"

namespace {

void *(*g_malloc)(size_t) = malloc;
void (*g_free)(void *) = free;
/* last error of the calling thread */
thread_local char g_err[512];
/* the device every call runs on (LIBPT_AMD_DEVICE, read once); the tune
 * cache is shared by all threads */
int g_dev = -1;
std::mutex g_mu;

void
set_err(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

double
now_s()
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#define HCHK(call, ret)                                                      \
	do {                                                                 \
		hipError_t e_ = (call);                                      \
		if (e_ != hipSuccess) {                                      \
			set_err("%s failed: %s", #call, hipGetErrorString(e_)); \
			return ret;                                          \
		}                                                            \
	} while (0)

/* choose the device once per process, then make it current on the
 * calling thread (HIP's current device is per thread, and plans, batches
 * and their streams may be used from any thread) */
bool
dev_init()
{
	int dev;
	{
		std::lock_guard<std::mutex> lk(g_mu);
		if (g_dev < 0) {
			int n = 0;
			if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
				set_err("no HIP device available");
				return false;
			}
			const char *e = getenv("LIBPT_AMD_DEVICE");
			const int d = e ? atoi(e) : 0;
			if (d < 0 || d >= n) {
				set_err("LIBPT_AMD_DEVICE=%d out of range (%d devices)",
				    d, n);
				return false;
			}
			g_dev = d;
		}
		dev = g_dev;
	}
	HCHK(hipSetDevice(dev), false);
	return true;
}

/* host scratch through the libpt_set_malloc hooks (src/misc.c:22-47) */
template <class T>
struct HostBuf {
	T *p = nullptr;
	size_t n = 0;
	bool alloc(size_t count)
	{
		p = (T *)g_malloc((count ? count : 1) * sizeof(T));
		n = count;
		if (!p)
			set_err("libpt malloc failed (%zu bytes)", count * sizeof(T));
		return p != nullptr;
	}
	~HostBuf()
	{
		if (p)
			g_free(p);
	}
};

/* element counts of the API arrays in pt.h order */
int
api_sizes(int call, int64_t oa, int64_t va, int64_t ob, int64_t vb,
    int64_t n[10])
{
	const bool ft = call == LIBPT_AMD_RFT || call == LIBPT_AMD_UFT;
	const bool r = call == LIBPT_AMD_RPT || call == LIBPT_AMD_RFT;
	int64_t ov, oovv, oovo, ovvv;
	if (r) {
		ov = oa * va;
		oovv = 2 * oa * oa * va * va;
		oovo = 2 * oa * oa * va * oa;
		ovvv = oa * va * (va * (va - 1) / 2) + oa * va * va * va;
	} else {
		ov = oa * va + ob * vb;
		oovv = oa * oa * va * va + 2 * oa * ob * va * vb + ob * ob * vb * vb;
		oovo = oa * oa * va * oa + oa * ob * va * ob + ob * ob * vb * ob +
		    ob * oa * vb * oa;
		ovvv = oa * va * (va * (va - 1) / 2) + oa * vb * va * vb +
		    ob * vb * (vb * (vb - 1) / 2) + ob * va * vb * va;
	}
	if (!ft) {
		int64_t s[7] = {ov, ov, ov, oovv, oovo, oovv, ovvv};
		memcpy(n, s, sizeof(s));
		return 7;
	}
	int64_t s[10] = {ov, ov, ov, oovv, oovv, oovv, oovo, ovvv, oovo, ovvv};
	memcpy(n, s, sizeof(s));
	return 10;
}

double
flop_same(double o, double v)
{
	return 3.0 * v * v * v * (v - 1) + 6.0 * o * v * v * v;
}

double
flop_mixed(double oa, double va, double ob, double vb)
{
	return 2 * va * va * va * vb + 4 * va * va * vb * vb +
	    2 * vb * va * va * (va - 1) + 6 * oa * va * va * vb +
	    4 * ob * va * va * vb;
}

/* blocking of the tile order (A, B, C tiles per block) and, when the
 * tuner recorded one, the tile geometry (waves per workgroup; 0: none) */
struct TileBlock {
	int ga = 8, gb = 8, gc = 2;
	int nw = 0;
};

/* LIBPT_AMD_TILE_BLOCK="ga,gb,gc" forces a blocking */
bool
tile_block_env(TileBlock &tb)
{
	const char *e = getenv("LIBPT_AMD_TILE_BLOCK");
	int a, b, c;
	if (e && sscanf(e, "%d,%d,%d", &a, &b, &c) == 3 && a > 0 && b > 0 &&
	    c > 0) {
		tb.ga = a;
		tb.gb = b;
		tb.gc = c;
		return true;
	}
	return false;
}

/*
 * Tuning cache: the tile geometry and blocking chosen by
 * libpt_amd_plan_tune per (form, call, oa, va, ob, vb), kept for the
 * process and, when LIBPT_AMD_TUNE_CACHE names a file, read from and
 * appended to it (one line per entry: form call oa va ob vb ga gb gc nw;
 * lines without nw keep the default geometry).  Lookups and stores hold
 * g_mu.
 */
struct TuneEntry {
	int form, call;
	int64_t oa, va, ob, vb;
	TileBlock tb;
};

std::vector<TuneEntry> g_tune;
bool g_tune_loaded = false;

void
tune_load()
{
	if (g_tune_loaded)
		return;
	g_tune_loaded = true;
	const char *path = getenv("LIBPT_AMD_TUNE_CACHE");
	FILE *f = path ? fopen(path, "r") : nullptr;
	if (!f)
		return;
	char line[256];
	while (fgets(line, sizeof(line), f)) {
		TuneEntry t;
		long long oa, va, ob, vb;
		const int n = sscanf(line, "%d %d %lld %lld %lld %lld %d %d %d %d",
		    &t.form, &t.call, &oa, &va, &ob, &vb, &t.tb.ga, &t.tb.gb,
		    &t.tb.gc, &t.tb.nw);
		if (n < 9)
			continue;
		if (n == 9 || (t.tb.nw != 4 && t.tb.nw != 8))
			t.tb.nw = 0;
		t.oa = oa;
		t.va = va;
		t.ob = ob;
		t.vb = vb;
		if (t.tb.ga > 0 && t.tb.gb > 0 && t.tb.gc > 0)
			g_tune.push_back(t);
	}
	fclose(f);
}

/* may_load: read LIBPT_AMD_TUNE_CACHE first if not yet read (the
 * allocation-free entry point only looks at entries already in memory) */
bool
tune_lookup(int form, int call, int64_t oa, int64_t va, int64_t ob,
    int64_t vb, TileBlock &tb, bool may_load)
{
	std::lock_guard<std::mutex> lk(g_mu);
	if (may_load)
		tune_load();
	/* the latest entry wins */
	for (size_t n = g_tune.size(); n-- > 0;) {
		const TuneEntry &t = g_tune[n];
		if (t.form == form && t.call == call && t.oa == oa && t.va == va &&
		    t.ob == ob && t.vb == vb) {
			tb = t.tb;
			return true;
		}
	}
	return false;
}

void
tune_store(int form, int call, int64_t oa, int64_t va, int64_t ob,
    int64_t vb, const TileBlock &tb)
{
	std::lock_guard<std::mutex> lk(g_mu);
	tune_load();
	g_tune.push_back({form, call, oa, va, ob, vb, tb});
	const char *path = getenv("LIBPT_AMD_TUNE_CACHE");
	FILE *f = path ? fopen(path, "a") : nullptr;
	if (f) {
		fprintf(f, "%d %d %lld %lld %lld %lld %d %d %d %d\n", form, call,
		    (long long)oa, (long long)va, (long long)ob, (long long)vb,
		    tb.ga, tb.gb, tb.gc, tb.nw);
		fclose(f);
	}
}

struct Case {
	TileBlock block;
	int same = 0, nA = 0, nC = 0, tcw = 16;  /* tile-list geometry */
	pt_launch L;
	pt_term_list terms;
	pt_term *d_terms = nullptr;
	int32_t *d_ijk = nullptr, *d_tiles = nullptr;
	double *d_partial = nullptr;
	int64_t n_ijk = 0;
	int n_tiles = 0;
	double factor = 1.0, flop_per_ijk = 0.0;
};

/* block pointers of one 4-index family: [aaaa, abab, bbbb, baba] */
struct Blocks {
	const char *p[4];
};

} /* namespace */

/* device memory of an allocation-free call: the plan takes its buffers
 * from [base, base + cap) in order, 256-byte aligned */
struct Arena {
	char *base = nullptr;
	size_t cap = 0, used = 0;
};

struct libpt_amd_plan {
	int call = 0, form = 0;
	int nw = 4;                 /* waves per workgroup of the tile kernel */
	int kgrp = 4;               /* ijk per item group of a launch: 4
	                               measured 10 % ahead of 1 for FP64 and
	                               4 % for split-BF16 at o=45, v=477 (2, 8
	                               and 16 within 1 % of 4) */
	bool vct = true;            /* VC packed a second time, transposed */
	bool fixed_geometry = false; /* a batch's plan: its launches are
	                                recorded, so the tuner keeps nw */
	bool unusable = false;      /* a geometry or order change failed half
	                               way: the tile lists are not whole */
	bool persist = false;       /* persistent grid (LIBPT_AMD_PERSISTENT) */
	uint32_t *d_ctr = nullptr;  /* its 8 item counters */
	int64_t oa = 0, va = 0, ob = 0, vb = 0;
	hipStream_t stream = nullptr;
	bool own_stream = true;     /* created (and destroyed) by the plan */
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	FixedVec<Case, 4> cases;
	/* device buffers from hipMalloc, freed with the plan; with an arena
	 * every buffer comes from it and nothing is allocated or freed */
	FixedVec<void *, 64> owned;
	Arena *arena = nullptr;
	double *d_out = nullptr;
	double scale[4] = {1, 1, 1, 1};
	double prep_s = 0;
	/* two W sets: save areas of one launch's items (wslots of
	 * pt_wsave_bytes each), allocated with the plan, or rebuilt by a
	 * geometry change; a batch's plans point at the batch's */
	void *wsave = nullptr;
	int64_t wslots = 0;
};

namespace {

/* waits for the plan's work, then frees what it owns (not the plan) */
void
plan_release(libpt_amd_plan *pl)
{
	/* a caller's stream may be the null stream */
	if (pl->stream || !pl->own_stream)
		(void)hipStreamSynchronize(pl->stream);
	for (void *p : pl->owned)
		(void)hipFree(p);
	pl->owned.clear();
	if (pl->ev0)
		(void)hipEventDestroy(pl->ev0);
	if (pl->ev1)
		(void)hipEventDestroy(pl->ev1);
	if (pl->stream && pl->own_stream)
		(void)hipStreamDestroy(pl->stream);
	pl->ev0 = pl->ev1 = nullptr;
	pl->stream = nullptr;
}

void
plan_free(libpt_amd_plan *pl)
{
	if (!pl)
		return;
	plan_release(pl);
	delete pl;
}

bool
dev_alloc(void **p, size_t bytes)
{
	hipError_t e = hipMalloc(p, bytes ? bytes : 256);
	if (e != hipSuccess) {
		set_err("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
		*p = nullptr;
		return false;
	}
	return true;
}

/* a device buffer of the plan: from its arena, else from hipMalloc
 * (freed with the plan) */
bool
plan_alloc(libpt_amd_plan *pl, void **p, size_t bytes)
{
	*p = nullptr;
	if (pl->arena) {
		Arena &a = *pl->arena;
		const size_t need = (bytes + 255) & ~(size_t)255;
		if (need > a.cap - a.used) {
			set_err("the call needs more than the %zu bytes of static "
			    "scratch of the allocation-free entry point (%zu in use, "
			    "%zu more requested)", a.cap, a.used, need);
			return false;
		}
		*p = a.base + a.used;
		a.used += need;
		return true;
	}
	if (pl->owned.size() >= 64) {
		set_err("too many device buffers in one plan");
		return false;
	}
	if (!dev_alloc(p, bytes))
		return false;
	pl->owned.push_back(*p);
	return true;
}

/* workgroups of a persistent grid: every CU's share for the geometry,
 * rounded up to a multiple of the 8 XCDs */
int
persist_grid(int nw)
{
	int ncu = 0;
	if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
	    g_dev) != hipSuccess || ncu <= 0)
		ncu = 256;
	const int g = ncu * (nw == 8 ? 1 : 2);
	return (g + 7) / 8 * 8;
}

/* frees one of the plan's hipMalloc buffers (no-op for null) */
void
plan_drop(libpt_amd_plan *pl, void *p)
{
	if (!p || pl->arena)
		return;
	for (int i = 0; i < pl->owned.size(); ++i)
		if (pl->owned[i] == p) {
			(void)hipFree(p);
			pl->owned[i] = pl->owned[pl->owned.size() - 1];
			pl->owned.n--;
			return;
		}
}

/* Host lists go to the device in chunks of N triples of ints through one
 * chunk buffer: host scratch from the libpt_set_malloc hooks (as the
 * reference takes its scratch, src/pt.c:208,223), or a stack array for
 * the allocation-free entry point. */
enum { LIST_CHUNK = 2048 };

struct ListBuf {
	int32_t stack_[3 * LIST_CHUNK];
	HostBuf<int32_t> heap_;
	int32_t *p = nullptr;
	bool get(const libpt_amd_plan *pl)
	{
		if (pl->arena) {
			p = stack_;
			return true;
		}
		if (!heap_.alloc(3 * LIST_CHUNK))
			return false;
		p = heap_.p;
		return true;
	}
};

struct ListUpload {
	int32_t *buf;
	int n = 0;
	int64_t done = 0;
	int32_t *dst;
	bool ok = true;
	ListUpload(int32_t *d, int32_t *b) : buf(b), dst(d) {}
	void put(int a, int b, int c)
	{
		buf[3 * n] = a;
		buf[3 * n + 1] = b;
		buf[3 * n + 2] = c;
		if (++n == LIST_CHUNK)
			flush();
	}
	bool flush()
	{
		if (n > 0 && ok) {
			hipError_t e = hipMemcpy(dst + 3 * done, buf,
			    3 * sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
			if (e != hipSuccess) {
				set_err("list upload failed: %s", hipGetErrorString(e));
				ok = false;
			}
		}
		done += n;
		n = 0;
		return ok;
	}
};

/* Tiles in blocks of GA x GB x GC (A, B, C): the workgroups that run at the
 * same time on one XCD take consecutive items, and a compact block lets
 * them share the Q slices of all three roles in L2 (Q[B][C] across A,
 * Q[A][C] across B, Q[A][B] across C).  Writes 3 ints per tile; returns
 * the tile count (the same for every blocking). */
template <class Emit>
int
tile_order(const Case &c, const TileBlock &tb, Emit emit)
{
	const int nA = c.nA, nC = c.nC, tcw = c.tcw;
	int nt = 0;
	for (int A0 = 0; A0 < nA; A0 += tb.ga)
	for (int B0 = 0; B0 < A0 + tb.ga && B0 < nA; B0 += tb.gb)
	for (int C0 = 0; C0 < nC; C0 += tb.gc)
		for (int A = A0; A < A0 + tb.ga && A < nA; ++A)
		for (int B = B0; B < B0 + tb.gb && B <= A; ++B)
		for (int Cb = C0; Cb < C0 + tb.gc && Cb < nC; ++Cb) {
			/* same spin needs some b > c: c_min < b_max */
			if (c.same && Cb * tcw >= B * PT_TB + PT_TB - 1)
				continue;
			emit(A, B, Cb);
			++nt;
		}
	return nt;
}

/* a new blocking for an existing case (same tiles, new order) */
bool
case_set_block(const libpt_amd_plan *pl, Case &c, const TileBlock &tb)
{
	if (c.n_tiles == 0)
		return true;
	ListBuf lb;
	if (!lb.get(pl))
		return false;
	int nt = 0;
	(void)tile_order(c, tb, [&](int, int, int) { ++nt; });
	if (nt != c.n_tiles) {
		set_err("tile order changed the tile count");
		return false;
	}
	ListUpload up(c.d_tiles, lb.p);
	(void)tile_order(c, tb, [&](int a, int b, int x) { up.put(a, b, x); });
	if (!up.flush())
		return false;
	c.block = tb;
	return true;
}

/* tile list of one spin case for c tiles of width tcw, and its item
 * energy slots (n_ijk x tiles, cleared).  A >= B; for the same-spin case
 * (a > b > c) only c tiles that reach below the top of the B block. */
bool
case_tiles(libpt_amd_plan *pl, Case &c, int64_t nc, int tcw)
{
	c.nC = (int)((nc + tcw - 1) / tcw);
	c.tcw = tcw;
	int nt = 0;
	(void)tile_order(c, c.block, [&](int, int, int) { ++nt; });
	c.n_tiles = nt;
	if (!plan_alloc(pl, (void **)&c.d_tiles, 3 * sizeof(int32_t) *
	    (size_t)nt) ||
	    !plan_alloc(pl, (void **)&c.d_partial,
	    sizeof(double) * (size_t)(c.n_ijk * nt)))
		return false;
	ListBuf lb;
	if (!lb.get(pl))
		return false;
	ListUpload ut(c.d_tiles, lb.p);
	(void)tile_order(c, c.block, [&](int a, int b, int x) { ut.put(a, b, x); });
	if (!ut.flush())
		return false;
	HCHK(hipMemset(c.d_partial, 0, sizeof(double) * (size_t)(c.n_ijk * nt)),
	    false);
	return true;
}

/* work list of one spin case: ijk items x (A, B, C) tiles */
bool
case_lists(libpt_amd_plan *pl, Case &c, int same, int64_t o1, int64_t o2,
    int64_t na, int64_t nc, int tcw)
{
	int64_t nijk = same ? o1 * (o1 - 1) * (o1 - 2) / 6 : o1 * (o1 - 1) / 2 * o2;
	c.same = same;
	c.nA = (int)((na + PT_TA - 1) / PT_TA);
	c.n_ijk = nijk;
	if (!plan_alloc(pl, (void **)&c.d_ijk, 3 * sizeof(int32_t) *
	    (size_t)nijk))
		return false;
	ListBuf lb;
	if (!lb.get(pl))
		return false;
	ListUpload ui(c.d_ijk, lb.p);
	for (int64_t i = 0; i < o1; ++i)
		for (int64_t j = i + 1; j < o1; ++j)
			for (int64_t k = same ? j + 1 : 0; k < (same ? o1 : o2); ++k)
				ui.put((int)i, (int)j, (int)k);
	if (!ui.flush())
		return false;
	return case_tiles(pl, c, nc, tcw);
}

bool
case_finish(libpt_amd_plan *pl, Case &c, int form, int kind, int nset,
    int64_t oa, int64_t va, int64_t ob, int64_t vb)
{
	pt_launch &L = c.L;
	L.nw = pl->nw;
	L.kgrp = pl->kgrp;
	L.ctr = pl->persist ? pl->d_ctr : nullptr;
	L.pgrid = persist_grid(pl->nw);
	L.kind = kind;
	L.nset = nset;
	L.oa = oa;
	L.va = va;
	L.ob = ob;
	L.vb = vb;
	L.na = (int32_t)va;
	L.nc = (int32_t)vb;
	L.tcw = pt_tcw(form, nset, pl->nw);
	for (int s = 0; s < nset; ++s) {
		if (kind == PT_AAA)
			pt_terms_same(c.terms, s, oa, va);
		else
			pt_terms_mixed(c.terms, s, oa, va, ob, vb, pl->vct);
	}
	/* Q slab offsets: API elements (K q_k per slab) -> bytes of the
	 * packed copy (pt_kgroups q_k groups per slab) */
	for (pt_term &t : c.terms) {
		const int64_t slab = (int64_t)t.K * t.q_k;
		const int64_t pslab = pt_kgroups(form, t.K) * t.q_k *
		    pt_group_bytes(form);
		for (int x = 0; x < 3; ++x)
			t.q_off[x] = t.q_off[x] / slab * pslab;
	}
	L.nterm1 = pt_terms_order(c.terms, &L.nterm_a);
	L.nterm2 = (int32_t)c.terms.size() / nset - L.nterm1;
	if (!plan_alloc(pl, (void **)&c.d_terms, sizeof(pt_term) *
	    (size_t)c.terms.size()))
		return false;
	HCHK(hipMemcpy(c.d_terms, c.terms.data(),
	    sizeof(pt_term) * (size_t)c.terms.size(), hipMemcpyHostToDevice),
	    false);
	L.terms = c.d_terms;
	if (!case_lists(pl, c, kind == PT_AAA, oa, ob, va, vb, L.tcw))
		return false;
	L.ijk = c.d_ijk;
	L.tiles = c.d_tiles;
	L.n_tiles = c.n_tiles;
	L.partial = c.d_partial;
	c.flop_per_ijk = (kind == PT_AAA ? flop_same((double)oa, (double)va) :
	    flop_mixed((double)oa, (double)va, (double)ob, (double)vb)) * nset;
	return true;
}

/* power-of-two scale bringing max |x| of a family to about 2^14 */
bool
family_scale(libpt_amd_plan *pl, const double *const *arr, const int64_t *n,
    int cnt, double *scale)
{
	HCHK(hipMemsetAsync(pl->d_out, 0, sizeof(double), pl->stream), false);
	for (int q = 0; q < cnt; ++q)
		if (n[q] > 0 && pt_launch_absmax(arr[q], n[q], pl->d_out,
		    pl->stream) != 0) {
			set_err("absmax launch failed");
			return false;
		}
	double mx = 0;
	HCHK(hipMemcpyAsync(&mx, pl->d_out, sizeof(double),
	    hipMemcpyDeviceToHost, pl->stream), false);
	HCHK(hipStreamSynchronize(pl->stream), false);
	*scale = mx > 0 ? ldexp(1.0, (int)floor(log2(16384.0 / mx))) : 1.0;
	return true;
}

} /* namespace */

namespace {

/* two W sets: the save areas of one launch's items, sized to the largest
 * spin case (at most pt_ft_chunk); no-op when present */
bool
plan_wsave(libpt_amd_plan *pl)
{
	if (pl->wsave)
		return true;
	int64_t most = 1;
	for (const Case &d : pl->cases)
		most = std::max(most, d.n_ijk * d.n_tiles);
	int64_t slots = std::min<int64_t>(pt_ft_chunk(pl->form, pl->nw), most);
	const int64_t wb = pt_wsave_bytes(pl->form, 2, pl->nw);
	if (pl->persist)
		/* one area per workgroup of the grid, whatever the item count */
		slots = persist_grid(pl->nw);
	else if (pl->arena) {
		/* what the scratch has left, in whole areas: launches then hold
		 * fewer items */
		const int64_t left = (int64_t)(pl->arena->cap - pl->arena->used) / wb;
		slots = std::min(slots, left > 0 ? left : 1);
	}
	if (!plan_alloc(pl, &pl->wsave, (size_t)slots * (size_t)wb))
		return false;
	pl->wslots = slots;
	return true;
}

/* A new tile geometry (waves per workgroup: 4 or 8) for a plan: the tile
 * lists, item slots and save areas are rebuilt for the new c tile width;
 * operands, terms and ijk lists stay.  Not for plans in a batch (their
 * launches are recorded in the batch's graph) or on an arena. */
bool
plan_set_geometry(libpt_amd_plan *pl, int nw)
{
	if (nw != 4 && nw != 8) {
		set_err("waves per workgroup must be 4 or 8, not %d", nw);
		return false;
	}
	if (nw == pl->nw)
		return true;
	if (pl->fixed_geometry || pl->arena) {
		set_err("the geometry of this plan is fixed");
		return false;
	}
	HCHK(hipStreamSynchronize(pl->stream), false);
	const bool had_wsave = pl->wsave != nullptr;
	/* from here a failure leaves the plan without tile lists */
	pl->unusable = true;
	plan_drop(pl, pl->wsave);
	pl->wsave = nullptr;
	pl->wslots = 0;
	pl->nw = nw;
	for (Case &c : pl->cases) {
		if (c.n_ijk == 0)
			continue;
		pt_launch &L = c.L;
		plan_drop(pl, c.d_tiles);
		plan_drop(pl, c.d_partial);
		c.d_tiles = nullptr;
		c.d_partial = nullptr;
		L.nw = nw;
		L.tcw = pt_tcw(pl->form, L.nset, nw);
		L.pgrid = persist_grid(nw);
		if (!case_tiles(pl, c, L.nc, L.tcw))
			return false;
		L.tiles = c.d_tiles;
		L.n_tiles = c.n_tiles;
		L.partial = c.d_partial;
	}
	if (had_wsave && !plan_wsave(pl))
		return false;
	HCHK(hipDeviceSynchronize(), false);
	pl->unusable = false;
	return true;
}

/* Stages the operands and work lists of one call into pl, which comes
 * with its arena (or none) and, when pl->own_stream is false, its stream.
 * own_wsave: allocate the (fT) save areas now, so that runs allocate
 * nothing (plans owned by a batch use the batch's).  On failure the
 * caller releases pl. */
bool
plan_init(libpt_amd_plan *pl, int call, int form, size_t oa_, size_t va_,
    size_t ob_, size_t vb_, const void *const *arr, bool own_wsave)
{
	if (call < 0 || call > 3 || form < 0 || form >= PT_NFORM) {
		set_err("bad call %d or form %d", call, form);
		return false;
	}
	const bool ft = call == LIBPT_AMD_RFT || call == LIBPT_AMD_UFT;
	const bool restricted = call == LIBPT_AMD_RPT || call == LIBPT_AMD_RFT;
	int64_t oa = (int64_t)oa_, va = (int64_t)va_;
	int64_t ob = restricted ? oa : (int64_t)ob_;
	int64_t vb = restricted ? va : (int64_t)vb_;
	pl->nw = pt_nw_default(form);
	if (const char *w = getenv("LIBPT_AMD_WAVES"))
		pl->nw = atoi(w) == 4 ? 4 : atoi(w) == 8 ? 8 : pl->nw;
	if (const char *g = getenv("LIBPT_AMD_KGROUP"))
		pl->kgrp = std::max(1, std::min(64, atoi(g)));
	if (const char *t = getenv("LIBPT_AMD_VC_TRANSPOSED"))
		pl->vct = atoi(t) != 0;
	/* the static scratch of the allocation-free call holds more systems
	 * without the second copy of the mixed ovvv block */
	if (pl->arena)
		pl->vct = false;
	if (const char *t = getenv("LIBPT_AMD_PERSISTENT"))
		pl->persist = atoi(t) != 0;
	pl->call = call;
	pl->form = form;
	pl->oa = oa;
	pl->va = va;
	pl->ob = ob;
	pl->vb = vb;
#define PFAIL(cond)                  \
	do {                         \
		if (!(cond))         \
			return false; \
	} while (0)
	if (pl->own_stream) {
		PFAIL(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) ==
		    hipSuccess || (set_err("hipStreamCreate failed"), false));
		PFAIL((hipEventCreate(&pl->ev0) == hipSuccess &&
		    hipEventCreate(&pl->ev1) == hipSuccess) ||
		    (set_err("hipEventCreate failed"), false));
	}
	PFAIL(plan_alloc(pl, (void **)&pl->d_out, 4 * sizeof(double)));
	PFAIL(plan_alloc(pl, (void **)&pl->d_ctr, 8 * sizeof(uint32_t)));

	int64_t n[10];
	const int narr = api_sizes(call, oa, va, ob, vb, n);
	(void)narr;
	const bool fp32 = form == PT_FORM_FP32;
	const bool split = form == PT_FORM_BF16X3 || form == PT_FORM_F16X3;
	const size_t esz = fp32 ? 4 : 8;     /* epilogue element size */
	const size_t osz = fp32 || split ? 4 : 8; /* operand element size */

	/* per W set, API array index of the amplitudes (S operand of ovvv
	 * terms, Q operand of oovo terms), the oovo integrals (S operand) and
	 * the ovvv integrals (Q operand):
	 * (T) {t2, oovo, ovvv}; (fT) set 0 = L {l2, i6, i7}, set 1 = R {t2, i2, i3} */
	const int nset_ = ft ? 2 : 1;
	int famT[2] = {3, 3}, famO[2] = {4, 4}, famV[2] = {6, 6};
	if (ft) {
		famT[0] = 4; famO[0] = 8; famV[0] = 9;
		famT[1] = 3; famO[1] = 6; famV[1] = 7;
	}
	const int KC = pt_kc(form), GB = pt_group_bytes(form);
	const int nblk = restricted ? 2 : 4;
	const double t_prep = now_s();

	/* block geometry [R slabs][K kappa][PL plane] and element offsets,
	 * src/pt.c:460-467 (restricted) and :487-512 (unrestricted) */
	struct Geo {
		int64_t R, K, PL, off;
	};
	auto geo = [&](int fam, int blk) {
		const int64_t PA = va * (va - 1) / 2, PB = vb * (vb - 1) / 2;
		Geo g[4];
		if (fam == 0) {        /* t2-like [o][o][v][v] */
			g[0] = {oa, oa, va * va, 0};
			g[1] = {oa, ob, va * vb, 0};
			g[2] = {ob, ob, vb * vb, 0};
			g[3] = {ob, oa, vb * va, 0};
		} else {               /* ovvv [o][v][v][v], same spin packed */
			g[0] = {oa, va, PA, 0};
			g[1] = {oa, vb, va * vb, 0};
			g[2] = {ob, vb, PB, 0};
			g[3] = {ob, va, vb * va, 0};
		}
		for (int b = 1; b < 4; ++b)
			g[b].off = g[b - 1].off + g[b - 1].R * g[b - 1].K * g[b - 1].PL;
		return g[blk];
	};

	/* F16X3 operand scales: powers of two bringing each family's max to
	 * about 2^14 (amplitudes, integrals) */
	for (int s = 0; s < nset_; ++s) {
		double sa = 1, si = 1;
		if (form == PT_FORM_F16X3) {
			const double *fa[1] = {(const double *)arr[famT[s]]};
			const double *fi[2] = {(const double *)arr[famO[s]],
			    (const double *)arr[famV[s]]};
			int64_t na_[1] = {n[famT[s]]}, ni_[2] = {n[famO[s]], n[famV[s]]};
			PFAIL(family_scale(pl, fa, na_, 1, &sa));
			PFAIL(family_scale(pl, fi, ni_, 2, &si));
		}
		pl->scale[2 * s] = sa;
		pl->scale[2 * s + 1] = si;
	}
	/* S operands stay in the API layout (hi|lo pairs for the split forms) */
	const void *sT[2] = {nullptr, nullptr}, *sO[2] = {nullptr, nullptr};
	for (int s = 0; s < nset_; ++s) {
		const int fam2[2] = {famT[s], famO[s]};
		const double sc2[2] = {pl->scale[2 * s], pl->scale[2 * s + 1]};
		const void *out2[2];
		for (int q = 0; q < 2; ++q) {
			if (!split) {
				out2[q] = arr[fam2[q]];
				continue;
			}
			void *dst;
			PFAIL(plan_alloc(pl, &dst, sizeof(uint32_t) *
			    (size_t)n[fam2[q]]));
			PFAIL(pt_launch_split((const double *)arr[fam2[q]],
			    (uint32_t *)dst, n[fam2[q]], form == PT_FORM_F16X3, sc2[q],
			    pl->stream) == 0 || (set_err("split launch failed"), false));
			out2[q] = dst;
		}
		sT[s] = out2[0];
		sO[s] = out2[1];
	}
	/* Q operands: packed copies of every amplitude and ovvv block */
	Blocks qT[2], qV[2];
	for (int s = 0; s < nset_; ++s) {
		for (int fam = 0; fam < 2; ++fam) {
			const int api = fam == 0 ? famT[s] : famV[s];
			const double sc = pl->scale[2 * s + (fam == 0 ? 0 : 1)];
			Blocks &dstb = fam == 0 ? qT[s] : qV[s];
			for (int b = 0; b < nblk; ++b) {
				const Geo g = geo(fam, b);
				const int64_t nkb = pt_kgroups(form, g.K);
				void *dst;
				PFAIL(plan_alloc(pl, &dst, (size_t)(g.R * nkb * g.PL) * GB));
				const char *src = (const char *)arr[api] + esz * (size_t)g.off;
				PFAIL(pt_launch_pack(src, dst, g.R, g.K, nkb, g.PL, KC, form,
				    sc, 0, pl->stream) == 0 ||
				    (set_err("pack launch failed"), false));
				dstb.p[b] = (const char *)dst;
			}
			if (restricted) {
				dstb.p[2] = dstb.p[0];
				dstb.p[3] = dstb.p[1];
			}
		}
	}
	/* the mixed ovvv blocks as the mixed-spin case reads its VC operand
	 * (VC[k, d, c, b] with (u, w) = (b or a, c)): packed again with the
	 * plane transposed, so that the 32 w of a DMA row are contiguous
	 * (abab plane [va][vb], baba [vb][va]) */
	Blocks qVt[2];
	for (int s = 0; s < nset_ && pl->vct; ++s) {
		const double sc = pl->scale[2 * s + 1];
		for (int b = 1; b < (restricted ? 2 : 4); b += 2) {
			const Geo g = geo(1, b);
			const int64_t nkb = pt_kgroups(form, g.K);
			void *dst;
			PFAIL(plan_alloc(pl, &dst, (size_t)(g.R * nkb * g.PL) * GB));
			const char *src = (const char *)arr[famV[s]] + esz *
			    (size_t)g.off;
			PFAIL(pt_launch_pack(src, dst, g.R, g.K, nkb, g.PL, KC, form, sc,
			    b == 1 ? va : vb, pl->stream) == 0 ||
			    (set_err("pack launch failed"), false));
			qVt[s].p[b] = (const char *)dst;
		}
		if (restricted)
			qVt[s].p[3] = qVt[s].p[1];
	}
	PFAIL(hipStreamSynchronize(pl->stream) == hipSuccess ||
	    (set_err("operand preparation failed"), false));
	pl->prep_s = now_s() - t_prep;

	/* block pointers (element offsets as in src/pt.c:460-467, 487-512) */
	auto blk4 = [&](const void *base, int fam, size_t el) {
		Blocks b;
		int64_t s[4];
		if (fam == 0) {
			s[0] = oa * oa * va * va; s[1] = oa * ob * va * vb;
			s[2] = ob * ob * vb * vb;
		} else if (fam == 1) {
			s[0] = oa * oa * va * oa; s[1] = oa * ob * va * ob;
			s[2] = ob * ob * vb * ob;
		} else {
			s[0] = oa * va * (va * (va - 1) / 2); s[1] = oa * vb * va * vb;
			s[2] = ob * vb * (vb * (vb - 1) / 2);
		}
		const char *p = (const char *)base;
		b.p[0] = p;
		b.p[1] = p + el * (size_t)s[0];
		b.p[2] = b.p[1] + el * (size_t)s[1];
		b.p[3] = b.p[2] + el * (size_t)s[2];
		if (restricted) { /* baba = abab, bbbb = aaaa */
			b.p[2] = b.p[0];
			b.p[3] = b.p[1];
		}
		return b;
	};
	auto ov2 = [&](const void *base, size_t el, const char **pa,
	    const char **pb) {
		*pa = (const char *)base;
		*pb = restricted ? *pa : *pa + el * (size_t)(oa * va);
	};

	const char *dA, *dB, *fA, *fB, *uA, *uB;
	ov2(arr[0], esz, &dA, &dB);
	ov2(arr[1], esz, &fA, &fB);
	ov2(arr[2], esz, &uA, &uB);
	/* epilogue T2-like (Z uses t2 for (T), l2 for (fT)) and oovv */
	Blocks zt = blk4(arr[ft ? 4 : 3], 0, esz);
	Blocks gv = blk4(arr[5], 0, esz);
	/* S operand blocks per W set */
	Blocks opT[2], opO[2];
	for (int s = 0; s < nset_; ++s) {
		opT[s] = blk4(sT[s], 0, osz);
		opO[s] = blk4(sO[s], 1, osz);
	}
	const int nset = ft ? 2 : 1;
	const double fac = restricted ? 2.0 : 1.0;

	/* spin cases: aaa(alpha), [aaa(beta)], aab(alpha pair), [aab(beta pair)]
	 * (src/pt.c:469-473 restricted; :514-529 unrestricted) */
	struct CaseSpec {
		int kind, spin;
	};
	std::vector<CaseSpec> specs;
	specs.push_back({PT_AAA, 0});
	if (!restricted)
		specs.push_back({PT_AAA, 1});
	specs.push_back({PT_AAB, 0});
	if (!restricted)
		specs.push_back({PT_AAB, 1});

	/* tile-order blocking: forced by the environment, else the tuning
	 * cache, else the default */
	TileBlock blk;
	if (!tile_block_env(blk))
		(void)tune_lookup(form, call, oa, va, ob, vb, blk, !pl->arena);
	if (blk.nw && !getenv("LIBPT_AMD_WAVES"))
		pl->nw = blk.nw;
	for (const CaseSpec &cs : specs) {
		Case c;
		c.block = blk;
		memset(&c.L, 0, sizeof(c.L));
		const int sp = cs.spin;
		/* block indices of the pair spin (same) and mixed blocks */
		const int same = sp == 0 ? 0 : 2;   /* aaaa or bbbb */
		const int mixB = sp == 0 ? 1 : 3;   /* abab or baba */
		const int mixC = sp == 0 ? 3 : 1;   /* baba or abab */
		int64_t o1 = sp == 0 ? oa : ob, v1 = sp == 0 ? va : vb;
		int64_t o2 = sp == 0 ? ob : oa, v2 = sp == 0 ? vb : va;
		for (int s = 0; s < nset; ++s) {
			const int sl = 12 * s;
			c.L.buf[sl + 0] = opT[s].p[same];
			c.L.buf[sl + 1] = opT[s].p[mixB];
			c.L.buf[sl + 2] = opT[s].p[mixC];
			c.L.buf[sl + 3] = opO[s].p[same];
			c.L.buf[sl + 4] = opO[s].p[mixB];
			c.L.buf[sl + 5] = opO[s].p[mixC];
			c.L.buf[sl + 6] = qV[s].p[same];
			c.L.buf[sl + 7] = qV[s].p[mixB];
			c.L.buf[sl + 8] = pl->vct ? qVt[s].p[mixC] : qV[s].p[mixC];
			c.L.buf[sl + 9] = qT[s].p[same];
			c.L.buf[sl + 10] = qT[s].p[mixB];
			c.L.buf[sl + 11] = qT[s].p[mixC];
			c.L.inv_scale[s] = 1.0 / (pl->scale[2 * s] * pl->scale[2 * s + 1]);
		}
		pt_epi &e = c.L.epi;
		e.d1 = sp == 0 ? dA : dB;
		e.d2 = sp == 0 ? dB : dA;
		e.u1 = sp == 0 ? uA : uB;
		e.u2 = sp == 0 ? uB : uA;
		e.f1 = sp == 0 ? fA : fB;
		e.f2 = sp == 0 ? fB : fA;
		e.GA = gv.p[same];
		e.GB = gv.p[mixB];
		e.ZA = zt.p[same];
		e.ZB = zt.p[mixB];
		c.factor = fac;
		bool empty;
		if (cs.kind == PT_AAA) {
			empty = o1 < 3 || v1 < 3;
			if (!empty)
				PFAIL(case_finish(pl, c, form, PT_AAA, nset, o1, v1, o1,
				    v1));
		} else {
			empty = o1 < 2 || v1 < 2 || o2 < 1 || v2 < 1;
			if (!empty)
				PFAIL(case_finish(pl, c, form, PT_AAB, nset, o1, v1, o2,
				    v2));
		}
		pl->cases.push_back(c);
	}
	if (ft && own_wsave)
		PFAIL(plan_wsave(pl));
	/* list uploads and slot clears ran on the null stream; the plan's
	 * stream does not wait for it */
	PFAIL(hipDeviceSynchronize() == hipSuccess ||
	    (set_err("hipDeviceSynchronize failed"), false));
#undef PFAIL
	return true;
}

libpt_amd_plan *
plan_create(int call, int form, size_t oa, size_t va, size_t ob, size_t vb,
    const void *const *arr, bool own_wsave)
{
	if (!dev_init())
		return nullptr;
	libpt_amd_plan *pl = new libpt_amd_plan;
	if (!plan_init(pl, call, form, oa, va, ob, vb, arr, own_wsave)) {
		plan_free(pl);
		return nullptr;
	}
	return pl;
}

} /* namespace */

namespace {

/*
 * The tile launches over items [item_b, item_e) of one spin case: grids of
 * at most 2^22 items, and with two W sets of at most wslots (one save area
 * per item).  With base.kgrp > 1, launches hold whole ijk (the kernel
 * takes them in groups of kgrp consecutive ijk); where a launch cannot
 * hold one whole ijk the items go in order.
 */
int
launch_items(int form, const pt_launch &base, int64_t item_b, int64_t item_e,
    void *wsave, int64_t wslots, hipStream_t st)
{
	const int64_t nt = base.n_tiles;
	int64_t chunk = base.nset == 2 && !base.ctr ? wslots : (int64_t)1 << 22;
	int kg = base.kgrp;
	if (kg > 1) {
		if (item_b % nt || item_e % nt || chunk < nt)
			kg = 1;
		else
			chunk = chunk / nt * nt;
	}
	for (int64_t it = item_b; it < item_e; it += chunk) {
		pt_launch L = base;
		L.item0 = it;
		L.n_chunk = std::min(chunk, item_e - it);
		L.kgrp = kg;
		L.wsave = wsave;
		L.nwsave = (int32_t)wslots;
		const int rc = pt_launch_tiles(form, &L, st);
		if (rc != 0)
			return rc;
	}
	return 0;
}

} /* namespace */

extern "C" libpt_amd_plan *
libpt_amd_plan_create(int call, int form, size_t oa, size_t va, size_t ob,
    size_t vb, const void *const *arr)
{
	return plan_create(call, form, oa, va, ob, vb, arr, true);
}

extern "C" void
libpt_amd_plan_destroy(libpt_amd_plan *pl)
{
	(void)dev_init();
	plan_free(pl);
}

extern "C" int
libpt_amd_plan_ncases(const libpt_amd_plan *pl)
{
	return pl ? (int)pl->cases.size() : 0;
}

extern "C" int64_t
libpt_amd_plan_nijk(const libpt_amd_plan *pl, int sc)
{
	return pl && sc >= 0 && sc < (int)pl->cases.size() ? pl->cases[sc].n_ijk : 0;
}

extern "C" int64_t
libpt_amd_plan_items(const libpt_amd_plan *pl, int sc)
{
	if (!pl || sc < 0 || sc >= (int)pl->cases.size())
		return 0;
	return pl->cases[sc].n_ijk * pl->cases[sc].n_tiles;
}

extern "C" int
libpt_amd_plan_geometry(const libpt_amd_plan *pl, int *out3)
{
	if (!pl || !out3)
		return -1;
	out3[0] = pl->nw;
	out3[1] = pt_tcw(pl->form, pl->call == LIBPT_AMD_RFT ||
	    pl->call == LIBPT_AMD_UFT ? 2 : 1, pl->nw);
	out3[2] = pl->kgrp;
	return 0;
}

extern "C" double
libpt_amd_plan_flop_per_ijk(const libpt_amd_plan *pl, int sc)
{
	return pl && sc >= 0 && sc < (int)pl->cases.size() ?
	    pl->cases[sc].flop_per_ijk : 0.0;
}

extern "C" double
libpt_amd_plan_factor(const libpt_amd_plan *pl, int sc)
{
	return pl && sc >= 0 && sc < (int)pl->cases.size() ?
	    pl->cases[sc].factor : 0.0;
}

extern "C" int
libpt_amd_plan_run(libpt_amd_plan *pl, int sc, int64_t b, int64_t e)
{
	if (!dev_init())
		return -1;
	if (!pl || sc < 0 || sc >= (int)pl->cases.size()) {
		set_err("bad plan or spin case");
		return -1;
	}
	if (pl->unusable) {
		set_err("the plan lost its tile lists in a failed geometry or order change");
		return -1;
	}
	Case &c = pl->cases[sc];
	if (c.n_ijk == 0)
		return 0;
	if (b < 0)
		b = 0;
	if (e > c.n_ijk)
		e = c.n_ijk;
	if (e <= b)
		return 0;
	if (c.L.nset == 2 && !plan_wsave(pl))
		return -1;
	const int rc = launch_items(pl->form, c.L, b * c.n_tiles,
	    e * c.n_tiles, pl->wsave, pl->wslots, pl->stream);
	if (rc != 0) {
		set_err("tile kernel launch failed: %s",
		    hipGetErrorString((hipError_t)rc));
		return -1;
	}
	return 0;
}

extern "C" int
libpt_amd_plan_sync(libpt_amd_plan *pl)
{
	if (!dev_init())
		return -1;
	if (!pl)
		return -1;
	HCHK(hipStreamSynchronize(pl->stream), -1);
	return 0;
}

extern "C" double
libpt_amd_plan_energy(libpt_amd_plan *pl)
{
	if (!dev_init())
		return NAN;
	if (!pl)
		return NAN;
	double e = 0.0;
	for (Case &c : pl->cases) {
		if (c.n_ijk == 0)
			continue;
		if (pt_launch_reduce(c.d_partial, c.n_ijk * c.n_tiles, pl->d_out,
		    pl->stream) != 0) {
			set_err("reduce launch failed");
			return NAN;
		}
		double x = 0;
		HCHK(hipMemcpyAsync(&x, pl->d_out, sizeof(double),
		    hipMemcpyDeviceToHost, pl->stream), NAN);
		HCHK(hipStreamSynchronize(pl->stream), NAN);
		e += c.factor * x;
	}
	return e;
}

extern "C" int
libpt_amd_plan_ijk_energies(libpt_amd_plan *pl, int sc, double *out)
{
	if (!dev_init())
		return -1;
	if (!pl || sc < 0 || sc >= (int)pl->cases.size())
		return -1;
	Case &c = pl->cases[sc];
	if (c.n_ijk == 0)
		return 0;
	double *d;
	if (!dev_alloc((void **)&d, sizeof(double) * (size_t)c.n_ijk))
		return -1;
	int rc = pt_launch_segsum(c.d_partial, c.n_ijk, c.n_tiles, d, pl->stream);
	if (rc == 0 && hipMemcpyAsync(out, d, sizeof(double) * (size_t)c.n_ijk,
	    hipMemcpyDeviceToHost, pl->stream) == hipSuccess &&
	    hipStreamSynchronize(pl->stream) == hipSuccess)
		rc = 0;
	else
		rc = -1;
	(void)hipFree(d);
	return rc;
}

extern "C" int
libpt_amd_plan_ijk_list(const libpt_amd_plan *pl, int sc, int32_t *out)
{
	if (!pl || sc < 0 || sc >= (int)pl->cases.size())
		return -1;
	const Case &c = pl->cases[sc];
	if (c.n_ijk == 0)
		return 0;
	HCHK(hipMemcpy(out, c.d_ijk, 3 * sizeof(int32_t) * (size_t)c.n_ijk,
	    hipMemcpyDeviceToHost), -1);
	return 0;
}

extern "C" int
libpt_amd_plan_timer_begin(libpt_amd_plan *pl)
{
	if (!dev_init())
		return -1;
	if (!pl)
		return -1;
	HCHK(hipEventRecord(pl->ev0, pl->stream), -1);
	return 0;
}

extern "C" double
libpt_amd_plan_timer_end_ms(libpt_amd_plan *pl)
{
	if (!dev_init())
		return NAN;
	if (!pl)
		return NAN;
	float ms = 0;
	HCHK(hipEventRecord(pl->ev1, pl->stream), NAN);
	HCHK(hipEventSynchronize(pl->ev1), NAN);
	HCHK(hipEventElapsedTime(&ms, pl->ev0, pl->ev1), NAN);
	return (double)ms;
}

extern "C" int
libpt_amd_plan_set_block(libpt_amd_plan *pl, int ga, int gb, int gc)
{
	if (!dev_init())
		return -1;
	if (!pl || ga <= 0 || gb <= 0 || gc <= 0) {
		set_err("bad plan or blocking");
		return -1;
	}
	TileBlock tb;
	tb.ga = ga;
	tb.gb = gb;
	tb.gc = gc;
	HCHK(hipStreamSynchronize(pl->stream), -1);
	for (Case &c : pl->cases)
		if (!case_set_block(pl, c, tb)) {
			/* a tile list may be half rewritten */
			pl->unusable = true;
			return -1;
		}
	return 0;
}

extern "C" int
libpt_amd_plan_set_geometry(libpt_amd_plan *pl, int nw)
{
	if (!dev_init())
		return -1;
	if (!pl) {
		set_err("null plan");
		return -1;
	}
	return plan_set_geometry(pl, nw) ? 0 : -1;
}

extern "C" int
libpt_amd_tune_record(int call, int form, size_t oa, size_t va, size_t ob,
    size_t vb, int ga, int gb, int gc, int waves)
{
	if (call < 0 || call > 3 || form < 0 || form >= PT_NFORM || ga <= 0 ||
	    gb <= 0 || gc <= 0 || (waves != 0 && waves != 4 && waves != 8)) {
		set_err("bad tuning record");
		return -1;
	}
	const bool restricted = call == LIBPT_AMD_RPT || call == LIBPT_AMD_RFT;
	TileBlock tb;
	tb.ga = ga;
	tb.gb = gb;
	tb.gc = gc;
	tb.nw = waves;
	tune_store(form, call, (int64_t)oa, (int64_t)va,
	    (int64_t)(restricted ? oa : ob), (int64_t)(restricted ? va : vb), tb);
	return 0;
}

/*
 * Autotuner of the tile geometry and order: for each tile geometry (4
 * waves, two workgroups per CU, with 16-wide c tiles in FP64 and 32-wide
 * in the other forms; 8 waves, one workgroup per CU, with 32-wide c
 * tiles; only the plan's own for a plan in a batch) and each
 * candidate blocking, time the first `sample` ijk triples of every spin
 * case (after one untimed pass per geometry), weight each case by its ijk
 * count, keep the fastest, and record it in the tuning cache.  The
 * energies of the sampled items are overwritten by the next run of the
 * plan.  out4 (optional): ga, gb, gc, waves.
 */
extern "C" int
libpt_amd_plan_tune(libpt_amd_plan *pl, int sample, int *out4)
{
	if (!dev_init())
		return -1;
	if (!pl) {
		set_err("null plan");
		return -1;
	}
	static const int cand[][3] = {
		{1, 1, 64}, {2, 2, 8}, {4, 4, 2}, {4, 4, 4}, {8, 4, 2},
		{8, 8, 1}, {8, 8, 2}, {16, 16, 1}, {4, 2, 8}, {2, 2, 16},
	};
	const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
	if (sample <= 0)
		sample = 8;
	const bool geo = !pl->fixed_geometry && !pl->arena &&
	    !getenv("LIBPT_AMD_WAVES");
	int nws[2] = {pl->nw, pl->nw == 8 ? 4 : 8};
	const int ngeo = geo ? 2 : 1;
	double best = 0;
	int bi = -1, bw = pl->nw;
	for (int gi = 0; gi < ngeo; ++gi) {
		if (!plan_set_geometry(pl, nws[gi]))
			return -1;
		for (int ci = -1; ci < ncand; ++ci) {
			/* ci = -1: one untimed warm pass with the first
			 * candidate */
			const int *cb = cand[ci < 0 ? 0 : ci];
			if (libpt_amd_plan_set_block(pl, cb[0], cb[1], cb[2]) != 0)
				return -1;
			double t = 0;
			for (int sc = 0; sc < (int)pl->cases.size(); ++sc) {
				const Case &c = pl->cases[sc];
				if (c.n_ijk == 0)
					continue;
				const int64_t m = c.n_ijk < sample ? c.n_ijk : sample;
				if (libpt_amd_plan_timer_begin(pl) != 0 ||
				    libpt_amd_plan_run(pl, sc, 0, m) != 0)
					return -1;
				const double ms = libpt_amd_plan_timer_end_ms(pl);
				if (isnan(ms))
					return -1;
				t += ms * (double)c.n_ijk / (double)m;
			}
			if (ci >= 0 && (bi < 0 || t < best)) {
				best = t;
				bi = ci;
				bw = nws[gi];
			}
		}
	}
	if (!plan_set_geometry(pl, bw) ||
	    libpt_amd_plan_set_block(pl, cand[bi][0], cand[bi][1], cand[bi][2]))
		return -1;
	TileBlock tb;
	tb.ga = cand[bi][0];
	tb.gb = cand[bi][1];
	tb.gc = cand[bi][2];
	tb.nw = geo ? bw : 0;
	tune_store(pl->form, pl->call, pl->oa, pl->va, pl->ob, pl->vb, tb);
	if (out4) {
		out4[0] = tb.ga;
		out4[1] = tb.gb;
		out4[2] = tb.gc;
		out4[3] = bw;
	}
	return 0;
}

/* ---------------------------------------------------------------------
 * Batches: many independent systems in one graph launch.  Each system is a
 * plan (operands staged and packed once, at creation); a segment is one
 * spin case of one system.  The graph holds the segments' tile launches in
 * one chain, ordered by cost per item (largest first, so the small items of
 * small systems fill the tail), and one batched reduction that leaves one
 * fixed-order sum per segment.
 */
struct libpt_amd_batch {
	int call = 0, form = 0, n = 0, nset = 1;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	std::vector<libpt_amd_plan *> plans;
	std::vector<int> seg_sys;
	std::vector<int> seg_of;      /* [system][spin case] -> segment or -1 */
	std::vector<double> seg_factor;
	pt_launch *d_segs = nullptr;
	int64_t *d_begin = nullptr;
	double *d_out = nullptr;
	void *wsave = nullptr;       /* two W sets: one save area per item of
	                                a launch, wslots of them */
	int64_t wslots = 0;
	int64_t total = 0;
	double flop = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
};

namespace {

void
batch_free(libpt_amd_batch *b)
{
	if (!b)
		return;
	if (b->stream)
		(void)hipStreamSynchronize(b->stream);
	if (b->exec)
		(void)hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void)hipGraphDestroy(b->graph);
	for (libpt_amd_plan *pl : b->plans)
		plan_free(pl);
	(void)hipFree(b->d_segs);
	(void)hipFree(b->d_begin);
	(void)hipFree(b->d_out);
	(void)hipFree(b->wsave);
	if (b->ev0)
		(void)hipEventDestroy(b->ev0);
	if (b->ev1)
		(void)hipEventDestroy(b->ev1);
	if (b->stream)
		(void)hipStreamDestroy(b->stream);
	delete b;
}

} /* namespace */

extern "C" libpt_amd_batch *
libpt_amd_batch_create(int call, int form, int n, const int64_t *dims,
    const void *const *dev_arrays)
{
	if (!dev_init())
		return nullptr;
	if (n <= 0 || !dims || !dev_arrays) {
		set_err("empty batch");
		return nullptr;
	}
	const bool ft = call == LIBPT_AMD_RFT || call == LIBPT_AMD_UFT;
	const int na = ft ? 10 : 7;
	libpt_amd_batch *b = new libpt_amd_batch;
	b->call = call;
	b->form = form;
	b->n = n;
	b->nset = ft ? 2 : 1;
#define BFAIL(cond)                   \
	do {                          \
		if (!(cond)) {        \
			batch_free(b); \
			return nullptr; \
		}                     \
	} while (0)
	BFAIL(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) ==
	    hipSuccess || (set_err("hipStreamCreate failed"), false));
	BFAIL((hipEventCreate(&b->ev0) == hipSuccess &&
	    hipEventCreate(&b->ev1) == hipSuccess) ||
	    (set_err("hipEventCreate failed"), false));

	struct Seg {
		pt_launch L;
		int sys, sc;
		double factor, cost;
		int64_t items;
	};
	std::vector<Seg> segs;
	for (int s = 0; s < n; ++s) {
		const int64_t *d = dims + 4 * s;
		libpt_amd_plan *pl = plan_create(call, form, (size_t)d[0],
		    (size_t)d[1], (size_t)d[2], (size_t)d[3], dev_arrays + na * s,
		    false);
		BFAIL(pl != nullptr);
		pl->fixed_geometry = true;
		b->plans.push_back(pl);
		for (int sc = 0; sc < (int)pl->cases.size(); ++sc) {
			Case &c = pl->cases[sc];
			b->flop += c.flop_per_ijk * (double)c.n_ijk;
			if (c.n_ijk == 0 || c.n_tiles == 0)
				continue;
			Seg g;
			g.sc = sc;
			g.L = c.L;
			g.L.item0 = 0;
			g.items = c.n_ijk * c.n_tiles;
			g.L.n_chunk = g.items;
			g.sys = s;
			g.factor = c.factor;
			g.cost = c.flop_per_ijk / c.n_tiles;
			segs.push_back(g);
		}
	}
	/* largest items first; ties keep system order */
	std::stable_sort(segs.begin(), segs.end(),
	    [](const Seg &x, const Seg &y) { return x.cost > y.cost; });
	const int nseg = (int)segs.size();
	std::vector<pt_launch> L(nseg);
	std::vector<int64_t> begin(nseg + 1, 0);
	for (int g = 0; g < nseg; ++g) {
		L[g] = segs[g].L;
		begin[g + 1] = begin[g] + segs[g].items;
		b->seg_sys.push_back(segs[g].sys);
		b->seg_factor.push_back(segs[g].factor);
	}
	b->seg_of.assign((size_t)n * 4, -1);
	for (int g = 0; g < nseg; ++g)
		b->seg_of[(size_t)segs[g].sys * 4 + segs[g].sc] = g;
	b->total = begin[nseg];
	BFAIL(dev_alloc((void **)&b->d_segs, sizeof(pt_launch) * (nseg + 1)));
	BFAIL(dev_alloc((void **)&b->d_begin, sizeof(int64_t) * (nseg + 1)));
	BFAIL(dev_alloc((void **)&b->d_out, sizeof(double) * (nseg + 1)));
	/* systems may differ in geometry (recorded by the tuner): save
	 * areas of the largest */
	int64_t wbytes = 1;
	for (const libpt_amd_plan *pl : b->plans)
		wbytes = std::max(wbytes, pt_wsave_bytes(form, 2, pl->nw));
	if (b->nset == 2 && b->total > 0) {
		int64_t most = 1;
		for (const Seg &g : segs)
			most = std::max(most, g.items);
		b->wslots = std::min<int64_t>(((int64_t)4 << 30) / wbytes, most);
		/* a persistent grid keeps one area per workgroup */
		for (const libpt_amd_plan *pl : b->plans)
			if (pl->persist)
				b->wslots = std::max<int64_t>(b->wslots,
				    persist_grid(pl->nw));
		BFAIL(dev_alloc(&b->wsave, (size_t)b->wslots * (size_t)wbytes));
	}
	if (nseg > 0) {
		BFAIL(hipMemcpy(b->d_segs, L.data(), sizeof(pt_launch) * nseg,
		    hipMemcpyHostToDevice) == hipSuccess ||
		    (set_err("segment upload failed"), false));
		BFAIL(hipMemcpy(b->d_begin, begin.data(),
		    sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice) ==
		    hipSuccess || (set_err("segment upload failed"), false));
	}

	/*
	 * The graph: each segment's items as launches of the spin case's own
	 * tile kernel (the one plans use), largest items first, in one chain,
	 * then the batched reduction.  Every launch of a system of this size
	 * range fills the GPU by itself: two concurrent branches (launches
	 * dealt by accumulated cost) were measured 2-3 % slower than the chain,
	 * which runs at the speed of the plans run one after another.
	 */
	BFAIL(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) ==
	    hipSuccess || (set_err("stream capture failed"), false));
	int rc = 0;
	for (int g = 0; g < nseg && rc == 0; ++g)
		rc = launch_items(form, L[g], 0, segs[g].items, b->wsave, b->wslots,
		    b->stream);
	if (rc == 0)
		rc = pt_launch_reduce_batch(b->d_segs, b->d_begin, nseg, b->d_out,
		    b->stream);
	hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
	BFAIL((rc == 0 && ec == hipSuccess) ||
	    (set_err("batch capture failed: %s",
	    hipGetErrorString(rc ? (hipError_t)rc : ec)), false));
	BFAIL(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) ==
	    hipSuccess || (set_err("graph instantiation failed"), false));
#undef BFAIL
	return b;
}

extern "C" void
libpt_amd_batch_destroy(libpt_amd_batch *b)
{
	(void)dev_init();
	batch_free(b);
}

extern "C" double
libpt_amd_batch_flop(const libpt_amd_batch *b)
{
	return b ? b->flop : 0.0;
}

extern "C" int64_t
libpt_amd_batch_items(const libpt_amd_batch *b)
{
	return b ? b->total : 0;
}

/*
 * Tile order of every system of a batch: a system whose sizes have a
 * recorded blocking takes it, the others are tuned on their own plan
 * (libpt_amd_plan_tune, which records its choice), so each distinct size is
 * timed once.  The blocking only rewrites a plan's tile list in place, so
 * the recorded graph reads the new order on its next replay.  Returns the
 * number of systems timed, or -1.
 */
extern "C" int
libpt_amd_batch_tune(libpt_amd_batch *b, int sample_ijk)
{
	if (!dev_init())
		return -1;
	if (!b) {
		set_err("null batch");
		return -1;
	}
	HCHK(hipStreamSynchronize(b->stream), -1);
	int timed = 0;
	for (libpt_amd_plan *pl : b->plans) {
		TileBlock tb;
		if (tune_lookup(pl->form, pl->call, pl->oa, pl->va, pl->ob, pl->vb,
		    tb, true)) {
			if (libpt_amd_plan_set_block(pl, tb.ga, tb.gb, tb.gc) != 0)
				return -1;
		} else {
			/* (fT): the tuning runs use the batch's save areas
			 * rather than allocating one set per plan */
			const bool lend = b->wsave && !pl->wsave;
			if (lend) {
				pl->wsave = b->wsave;
				pl->wslots = b->wslots;
			}
			const int rc = libpt_amd_plan_tune(pl, sample_ijk, nullptr);
			if (lend) {
				pl->wsave = nullptr;
				pl->wslots = 0;
			}
			if (rc != 0)
				return -1;
			++timed;
		}
	}
	return timed;
}

/* GPU time of the last libpt_amd_batch_run, ms (synchronises) */
extern "C" double
libpt_amd_batch_last_ms(libpt_amd_batch *b)
{
	if (!dev_init())
		return NAN;
	if (!b)
		return NAN;
	float ms = 0;
	HCHK(hipEventSynchronize(b->ev1), NAN);
	HCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1), NAN);
	return (double)ms;
}

extern "C" int
libpt_amd_batch_run(libpt_amd_batch *b)
{
	if (!dev_init())
		return -1;
	if (!b) {
		set_err("null batch");
		return -1;
	}
	for (const libpt_amd_plan *pl : b->plans)
		if (pl->unusable) {
			set_err("a system's tile list was left half rewritten by a "
			    "failed tuning step");
			return -1;
		}
	HCHK(hipEventRecord(b->ev0, b->stream), -1);
	HCHK(hipGraphLaunch(b->exec, b->stream), -1);
	HCHK(hipEventRecord(b->ev1, b->stream), -1);
	return 0;
}

extern "C" int
libpt_amd_batch_energies(libpt_amd_batch *b, double *out)
{
	if (!dev_init())
		return -1;
	if (!b || !out) {
		set_err("null batch or output");
		return -1;
	}
	const int nseg = (int)b->seg_sys.size();
	std::vector<double> e(nseg > 0 ? nseg : 1);
	if (nseg > 0)
		HCHK(hipMemcpyAsync(e.data(), b->d_out, sizeof(double) * nseg,
		    hipMemcpyDeviceToHost, b->stream), -1);
	HCHK(hipStreamSynchronize(b->stream), -1);
	/* spin cases of one system in plan order, as libpt_amd_plan_energy */
	for (int s = 0; s < b->n; ++s) {
		double x = 0.0;
		const int ncase = (int)b->plans[s]->cases.size();
		for (int ci = 0; ci < ncase; ++ci) {
			const int g = b->seg_of[(size_t)s * 4 + ci];
			if (g >= 0)
				x += b->seg_factor[g] * e[g];
		}
		out[s] = x;
	}
	return 0;
}

extern "C" void *
libpt_amd_dev_alloc(size_t bytes)
{
	if (!dev_init())
		return nullptr;
	void *p;
	return dev_alloc(&p, bytes) ? p : nullptr;
}

extern "C" void
libpt_amd_dev_free(void *p)
{
	if (p)
		(void)hipFree(p);
}

extern "C" int
libpt_amd_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
	if (!dev_init())
		return -1;
	HCHK(hipMemcpy(dst, src, bytes, (hipMemcpyKind)kind), -1);
	return 0;
}

extern "C" int64_t
libpt_amd_synth_size(const void *spec, int tensor)
{
	return pt_synth_size((const pt_synth_spec *)spec, tensor);
}

extern "C" int
libpt_amd_synth(const void *spec_, int tensor, void *dst, int on_device,
    int fp32)
{
	const pt_synth_spec *spec = (const pt_synth_spec *)spec_;
	int64_t n = pt_synth_size(spec, tensor);
	if (on_device) {
		if (!dev_init())
			return -1;
		if (pt_launch_synth(spec, tensor, dst, n, fp32, nullptr) != 0) {
			set_err("synth launch failed");
			return -1;
		}
		HCHK(hipDeviceSynchronize(), -1);
		return 0;
	}
	for (int64_t t = 0; t < n; ++t) {
		double x = pt_synth_elem(spec, tensor, t);
		if (fp32)
			((float *)dst)[t] = (float)x;
		else
			((double *)dst)[t] = x;
	}
	return 0;
}

extern "C" int
libpt_amd_synth_range(const void *spec_, int tensor, void *dst, int64_t begin,
    int64_t count, int fp32)
{
	const pt_synth_spec *spec = (const pt_synth_spec *)spec_;
	const int64_t n = pt_synth_size(spec, tensor);
	if (begin < 0 || count < 0 || begin > n || count > n - begin) {
		set_err("synth range [%lld, +%lld) outside tensor of %lld",
		    (long long)begin, (long long)count, (long long)n);
		return -1;
	}
	for (int64_t t = 0; t < count; ++t) {
		double x = pt_synth_elem(spec, tensor, begin + t);
		if (fp32)
			((float *)dst)[t] = (float)x;
		else
			((double *)dst)[t] = x;
	}
	return 0;
}

extern "C" int
libpt_amd_copy_kernel(void *dst, const void *src, int64_t n)
{
	if (!dev_init())
		return -1;
	if (pt_launch_copy((const double *)src, (double *)dst, n, nullptr) != 0) {
		set_err("copy launch failed");
		return -1;
	}
	HCHK(hipDeviceSynchronize(), -1);
	return 0;
}

extern "C" const char *
libpt_amd_last_error(void)
{
	return g_err;
}

extern "C" int
pt_device_init(void)
{
	return dev_init() ? g_dev : -1;
}

extern "C" void
pt_set_error(const char *msg)
{
	if (msg)
		set_err("%s", msg);
	else
		g_err[0] = 0;
}

/*
 * The body of libpt_amd_energy_noalloc (pt_noalloc.hip): one call on
 * device-resident inputs whose every device buffer -- packed operands,
 * split operands, term tables, work lists, per-item energies, (fT) save
 * areas -- comes from the static scratch [scratch, scratch + cap); the plan
 * itself lives on this stack frame and its lists go up through a stack
 * buffer.  Runs on the caller's stream and synchronises it.
 */
extern "C" double
pt_noalloc_energy(int call, int form, size_t oa, size_t va, size_t ob,
    size_t vb, const void *const *dev_arrays, void *scratch, size_t cap,
    void *stream)
{
	if (!dev_init())
		return NAN;
	Arena ar;
	ar.base = (char *)scratch;
	ar.cap = cap;
	libpt_amd_plan pl;
	pl.arena = &ar;
	pl.own_stream = false;
	pl.stream = (hipStream_t)stream;
	double e = NAN;
	if (plan_init(&pl, call, form, oa, va, ob, vb, dev_arrays, true)) {
		bool ok = true;
		for (int sc = 0; sc < pl.cases.size() && ok; ++sc)
			ok = libpt_amd_plan_run(&pl, sc, 0, pl.cases[sc].n_ijk) == 0;
		if (ok)
			e = libpt_amd_plan_energy(&pl);
	}
	plan_release(&pl);
	return e;
}

namespace {

/*
 * Host-to-device copy of the caller's pageable arrays.  Small inputs go
 * with hipMemcpy.  Large ones are cut into chunks of H2D_CHUNK bytes that
 * H2D_WORKERS host threads take in turn: a worker copies its chunk into
 * one of its two pinned buffers and hands it to its own stream, so that
 * the host copies of several threads and the PCIe transfers overlap.
 */
enum { H2D_WORKERS = 8 };
constexpr size_t H2D_CHUNK = (size_t)32 << 20;
constexpr size_t H2D_PINNED_FROM = (size_t)256 << 20;

bool
h2d_copy(void *const *dst, const void *const *src, const size_t *bytes,
    int n)
{
	size_t total = 0;
	for (int q = 0; q < n; ++q)
		total += bytes[q];
	if (total < H2D_PINNED_FROM) {
		for (int q = 0; q < n; ++q)
			if (bytes[q] > 0)
				HCHK(hipMemcpy(dst[q], src[q], bytes[q],
				    hipMemcpyHostToDevice), false);
		return true;
	}
	/* chunk list: (array, offset) in array order */
	std::vector<std::pair<int, size_t>> chunks;
	for (int q = 0; q < n; ++q)
		for (size_t o = 0; o < bytes[q]; o += H2D_CHUNK)
			chunks.push_back({q, o});
	const unsigned hw = std::thread::hardware_concurrency();
	const int nw = (int)std::max(1u, std::min<unsigned>(H2D_WORKERS,
	    hw > 2 ? hw / 2 : 1));
	std::atomic<size_t> next(0);
	std::atomic<bool> failed(false);
	int dev = 0;
	HCHK(hipGetDevice(&dev), false);
	std::mutex emu;
	std::string emsg;
	auto fail = [&](const char *what, hipError_t e) {
		std::lock_guard<std::mutex> lk(emu);
		if (!failed.exchange(true))
			emsg = std::string(what) + ": " + hipGetErrorString(e);
	};
	auto worker = [&]() {
		hipStream_t st = nullptr;
		void *buf[2] = {nullptr, nullptr};
		hipEvent_t ev[2] = {nullptr, nullptr};
		hipError_t e = hipSetDevice(dev);
		if (e == hipSuccess)
			e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
		for (int b = 0; b < 2 && e == hipSuccess; ++b) {
			e = hipHostMalloc(&buf[b], H2D_CHUNK, hipHostMallocDefault);
			if (e == hipSuccess)
				e = hipEventCreateWithFlags(&ev[b],
				    hipEventDisableTiming);
		}
		if (e != hipSuccess)
			fail("staging set-up failed", e);
		for (size_t k = 0, use = 0; e == hipSuccess && !failed; ++use) {
			k = next.fetch_add(1);
			if (k >= chunks.size())
				break;
			const int b = (int)(use & 1);
			const int q = chunks[k].first;
			const size_t o = chunks[k].second;
			const size_t len = std::min(H2D_CHUNK, bytes[q] - o);
			/* the buffer's previous transfer must be done */
			if ((e = hipEventSynchronize(ev[b])) != hipSuccess)
				break;
			memcpy(buf[b], (const char *)src[q] + o, len);
			e = hipMemcpyAsync((char *)dst[q] + o, buf[b], len,
			    hipMemcpyHostToDevice, st);
			if (e == hipSuccess)
				e = hipEventRecord(ev[b], st);
		}
		if (e != hipSuccess)
			fail("host to device copy failed", e);
		if (st)
			(void)hipStreamSynchronize(st);
		for (int b = 0; b < 2; ++b) {
			if (ev[b])
				(void)hipEventDestroy(ev[b]);
			if (buf[b])
				(void)hipHostFree(buf[b]);
		}
		if (st)
			(void)hipStreamDestroy(st);
	};
	std::vector<std::thread> th;
	try {
		for (int w = 0; w < nw; ++w)
			th.emplace_back(worker);
	} catch (const std::exception &) {
		/* the workers started take every chunk; with none, copy here */
		if (th.empty())
			worker();
	}
	for (std::thread &t : th)
		t.join();
	if (failed) {
		set_err("%s", emsg.c_str());
		return false;
	}
	return true;
}

/* Host arrays -> HBM -> plan -> energy.  Returns NaN on failure. */
double
host_call(int call, int form, size_t oa, size_t va, size_t ob, size_t vb,
    const void *const *arrays, libpt_amd_stats *st)
{
	double t0 = now_s();
	if (!dev_init())
		return NAN;
	const bool restricted = call == LIBPT_AMD_RPT || call == LIBPT_AMD_RFT;
	if (restricted) {
		ob = oa;
		vb = va;
	}
	int64_t n[10];
	const int narr = api_sizes(call, (int64_t)oa, (int64_t)va, (int64_t)ob,
	    (int64_t)vb, n);
	const size_t el = form == PT_FORM_FP32 ? 4 : 8;
	void *d[10] = {nullptr};
	double e = NAN;
	libpt_amd_plan *pl = nullptr;
	double t1 = 0, t2 = 0;
	bool ok = true;
	size_t nbytes[10];
	for (int q = 0; q < narr && ok; ++q) {
		nbytes[q] = el * (size_t)n[q];
		ok = dev_alloc(&d[q], nbytes[q]);
	}
	if (ok)
		ok = h2d_copy(d, arrays, nbytes, narr);
	t1 = now_s();
	if (ok)
		pl = libpt_amd_plan_create(call, form, oa, va, ob, vb, d);
	if (pl) {
		t2 = now_s();
		/* LIBPT_AMD_PROGRESS set: each spin case in 16 parts, with a line
		 * on stderr after each (a synchronisation per part) */
		const bool progress = getenv("LIBPT_AMD_PROGRESS") != nullptr;
		const int64_t parts = progress ? 16 : 1;
		for (int sc = 0; sc < (int)pl->cases.size() && ok; ++sc) {
			const int64_t nijk = pl->cases[sc].n_ijk;
			for (int64_t p = 0; p < parts && ok; ++p) {
				const int64_t b = nijk * p / parts;
				const int64_t en = nijk * (p + 1) / parts;
				ok = libpt_amd_plan_run(pl, sc, b, en) == 0;
				if (ok && progress) {
					ok = libpt_amd_plan_sync(pl) == 0;
					fprintf(stderr, "[libpt_amd] spin case %d: ijk %lld of "
					    "%lld done, %.1f s\n", sc, (long long)en,
					    (long long)nijk, now_s() - t2);
				}
			}
		}
		if (ok)
			e = libpt_amd_plan_energy(pl);
	}
	double t3 = now_s();
	if (st) {
		memset(st, 0, sizeof(*st));
		st->h2d_s = t1 - t0;
		if (pl) {
			st->prep_s = pl->prep_s;
			st->kernel_s = t3 - t2;
			for (Case &c : pl->cases)
				st->flop += c.flop_per_ijk * (double)c.n_ijk;
			for (int s = 0; s < 4; ++s)
				st->scale[s] = pl->scale[s];
		}
	}
	plan_free(pl);
	for (int q = 0; q < narr; ++q)
		(void)hipFree(d[q]);
	if (st)
		st->wall_s = now_s() - t0;
	return e;
}

/* pt.h semantics: a failure prints and exits with status 1, as the
 * reference's err(1, ...) (src/pt.c:208-209, 223-224). */
double
dropin(int call, int form, size_t oa, size_t va, size_t ob, size_t vb,
    const void *const *arrays)
{
	g_err[0] = 0;
	double e = host_call(call, form, oa, va, ob, vb, arrays, nullptr);
	if (isnan(e) && g_err[0]) {
		fprintf(stderr, "libpt: %s\n", g_err);
		exit(1);
	}
	return e;
}

} /* namespace */

extern "C" double
libpt_amd_energy(int call, int form, size_t oa, size_t va, size_t ob,
    size_t vb, const void *const *arrays, struct libpt_amd_stats *stats)
{
	g_err[0] = 0;
	return host_call(call, form, oa, va, ob, vb, arrays, stats);
}

/* ------------------------------------------------------ pt.h drop-in */

extern "C" void
libpt_print_banner(void)
{
	printf("libpt-compatible perturbative triples for AMD Instinct MI355X\n");
	printf("CCSD(T) and EOM-CCSD(fT) on CDNA4 matrix cores\n");
}

extern "C" void
libpt_set_malloc(void *(*fn)(size_t))
{
	g_malloc = fn ? fn : malloc;
}

extern "C" void
libpt_set_free(void (*fn)(void *))
{
	g_free = fn ? fn : free;
}

extern "C" double
libpt_rpt(size_t oa, size_t va, const double *d_ov, const double *f_ov,
    const double *t1, const double *t2, const double *i_oovo,
    const double *i_oovv, const double *i_ovvv)
{
	const void *a[7] = {d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv};
	return dropin(LIBPT_AMD_RPT, PT_FORM_FP64, oa, va, oa, va, a);
}

extern "C" double
libpt_rpt_mp(size_t oa, size_t va, const float *d_ov, const float *f_ov,
    const float *t1, const float *t2, const float *i_oovo,
    const float *i_oovv, const float *i_ovvv)
{
	const void *a[7] = {d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv};
	return dropin(LIBPT_AMD_RPT, PT_FORM_FP32, oa, va, oa, va, a);
}

extern "C" double
libpt_upt(size_t oa, size_t va, size_t ob, size_t vb, const double *d_ov,
    const double *f_ov, const double *t1, const double *t2,
    const double *i_oovo, const double *i_oovv, const double *i_ovvv)
{
	const void *a[7] = {d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv};
	return dropin(LIBPT_AMD_UPT, PT_FORM_FP64, oa, va, ob, vb, a);
}

extern "C" double
libpt_upt_mp(size_t oa, size_t va, size_t ob, size_t vb, const float *d_ov,
    const float *f_ov, const float *t1, const float *t2, const float *i_oovo,
    const float *i_oovv, const float *i_ovvv)
{
	const void *a[7] = {d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv};
	return dropin(LIBPT_AMD_UPT, PT_FORM_FP32, oa, va, ob, vb, a);
}

extern "C" double
libpt_rft(size_t oa, size_t va, const double *d_ov, const double *f2_ov,
    const double *l1, const double *t2, const double *l2, const double *i_oovv,
    const double *i2_t2f2_oovo, const double *i3_ovvv, const double *i6_oovo,
    const double *i7_ovvv)
{
	const void *a[10] = {d_ov, f2_ov, l1, t2, l2, i_oovv, i2_t2f2_oovo,
	    i3_ovvv, i6_oovo, i7_ovvv};
	return dropin(LIBPT_AMD_RFT, PT_FORM_FP64, oa, va, oa, va, a);
}

extern "C" double
libpt_rft_mp(size_t oa, size_t va, const float *d_ov, const float *f2_ov,
    const float *l1, const float *t2, const float *l2, const float *i_oovv,
    const float *i2_t2f2_oovo, const float *i3_ovvv, const float *i6_oovo,
    const float *i7_ovvv)
{
	const void *a[10] = {d_ov, f2_ov, l1, t2, l2, i_oovv, i2_t2f2_oovo,
	    i3_ovvv, i6_oovo, i7_ovvv};
	return dropin(LIBPT_AMD_RFT, PT_FORM_FP32, oa, va, oa, va, a);
}

extern "C" double
libpt_uft(size_t oa, size_t va, size_t ob, size_t vb, const double *d_ov,
    const double *f2_ov, const double *l1, const double *t2, const double *l2,
    const double *i_oovv, const double *i2_t2f2_oovo, const double *i3_ovvv,
    const double *i6_oovo, const double *i7_ovvv)
{
	const void *a[10] = {d_ov, f2_ov, l1, t2, l2, i_oovv, i2_t2f2_oovo,
	    i3_ovvv, i6_oovo, i7_ovvv};
	return dropin(LIBPT_AMD_UFT, PT_FORM_FP64, oa, va, ob, vb, a);
}

extern "C" double
libpt_uft_mp(size_t oa, size_t va, size_t ob, size_t vb, const float *d_ov,
    const float *f2_ov, const float *l1, const float *t2, const float *l2,
    const float *i_oovv, const float *i2_t2f2_oovo, const float *i3_ovvv,
    const float *i6_oovo, const float *i7_ovvv)
{
	const void *a[10] = {d_ov, f2_ov, l1, t2, l2, i_oovv, i2_t2f2_oovo,
	    i3_ovvv, i6_oovo, i7_ovvv};
	return dropin(LIBPT_AMD_UFT, PT_FORM_FP32, oa, va, ob, vb, a);
}
