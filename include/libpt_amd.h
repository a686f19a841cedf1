/*
 * Extension API of the MI355X perturbative-triples engine.  These entry
 * points have no counterpart in the reference libpt (src/pt.h); they add
 * what the drop-in API of pt.h cannot express:
 *   - a choice of arithmetic form (FP64, FP32, split BF16 / FP16),
 *   - inputs that already live in GPU memory (no host copy),
 *   - running a subset of the ijk work list and reading per-ijk energies,
 *   - the tile geometry and order of the kernel, and their autotuner,
 *   - batches of many small systems in one recorded graph,
 *   - a call that allocates no device memory,
 *   - the deterministic synthetic-input generator used by tests and bench.
 *
 * Array arguments are always passed in pt.h order:
 *   (T)  : d_ov, f_ov, t1, t2, i_oovo, i_oovv, i_ovvv
 *   (fT) : d_ov, f2_ov, l1, t2, l2, i_oovv, i2_t2f2_oovo, i3_ovvv,
 *          i6_oovo, i7_ovvv
 * with the layouts of pt.h.  Element type: double for the FP64, BF16X3
 * and F16X3 forms, float for the FP32 form.
 */
#ifndef LIBPT_AMD_H
#define LIBPT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum libpt_amd_form {
	LIBPT_AMD_FP64 = 0,   /* FP64 MFMA (libpt_rpt & co.) */
	LIBPT_AMD_FP32 = 1,   /* FP32 MFMA (libpt_*_mp) */
	LIBPT_AMD_BF16X3 = 2, /* hi+lo BF16 products, FP32 accumulation */
	LIBPT_AMD_F16X3 = 3   /* hi+lo FP16 products with power-of-two scaling */
};

enum libpt_amd_call {
	LIBPT_AMD_RPT = 0, LIBPT_AMD_UPT = 1, LIBPT_AMD_RFT = 2, LIBPT_AMD_UFT = 3
};

struct libpt_amd_stats {
	double wall_s;    /* whole call */
	double h2d_s;     /* host -> device copies */
	double prep_s;    /* operand preparation (split, scaling) */
	double kernel_s;  /* tile kernels and reductions */
	double flop;      /* reference-formulation GEMM flops of the call */
	double scale[4];  /* operand scales (F16X3): amp/int of W sets 0, 1 */
};

/* One call with host arrays in a chosen form.  Returns the energy with
 * the same meaning as the pt.h function of `call`; NaN on error. */
double libpt_amd_energy(int call, int form, size_t oa, size_t va, size_t ob,
    size_t vb, const void *const *arrays, struct libpt_amd_stats *stats);

/* Allocation-free variant of libpt_amd_energy.  No counterpart in the
 * reference libpt: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The inputs are device pointers in pt.h order
 * (dev_arrays, read only); the work runs on `stream` (a hipStream_t, NULL
 * for the null stream), which the call synchronises; the energy is
 * returned, NaN on error (libpt_amd_last_error says why).  The call
 * allocates nothing: every buffer it needs (packed and split operands,
 * work lists, per-item energies, (fT) save areas) comes from one static
 * device array of LIBPT_AMD_NOALLOC_SCRATCH bytes per GPU, which each GPU
 * the library runs on holds until the process ends.  A call whose
 * buffers do not fit is rejected with an error (FP64 restricted (T) fits
 * up to about o=10, v=120).  Calls on one GPU share the array: overlapping
 * calls would race on it, so the library runs them one at a time. */
#define LIBPT_AMD_NOALLOC_SCRATCH ((size_t)256 << 20)
double libpt_amd_energy_noalloc(int call, int form, size_t oa, size_t va,
    size_t ob, size_t vb, const void *const *dev_arrays, void *stream);

/* ---- plans over device-resident inputs ---- */
typedef struct libpt_amd_plan libpt_amd_plan;

/* dev_arrays: device pointers in pt.h order; the plan keeps pointers
 * (not copies) for FP64/FP32 and allocates split operands for the split
 * forms.  NULL on error. */
libpt_amd_plan *libpt_amd_plan_create(int call, int form, size_t oa,
    size_t va, size_t ob, size_t vb, const void *const *dev_arrays);
void libpt_amd_plan_destroy(libpt_amd_plan *plan);

/* spin cases of the call (rpt/rft: 2, upt/uft: 4) and their sizes */
int libpt_amd_plan_ncases(const libpt_amd_plan *plan);
int64_t libpt_amd_plan_nijk(const libpt_amd_plan *plan, int sc);
int64_t libpt_amd_plan_items(const libpt_amd_plan *plan, int sc);
double libpt_amd_plan_flop_per_ijk(const libpt_amd_plan *plan, int sc);
/* multiplier of spin case sc in the call's energy (2 for restricted) */
double libpt_amd_plan_factor(const libpt_amd_plan *plan, int sc);
/* tile-kernel geometry of the plan: out3 = (waves per workgroup, c tile
 * width, ijk per item group).  Default: 4 waves (32 x 32 x 16 tiles for
 * FP64, 32 x 32 x 32 for FP32 and the split forms), unless the tuning
 * cache holds a geometry for the plan's sizes; items taken in groups of 4
 * consecutive ijk (tile-major within a group, so that the items of one
 * tile that share the slabs of i and j run together).  LIBPT_AMD_WAVES=4|8
 * and LIBPT_AMD_KGROUP=n (1 to 64) override them for new plans;
 * LIBPT_AMD_PERSISTENT=1 runs new plans' launches on a persistent grid
 * (one workgroup per CU share, items from per-XCD counters, one (fT)
 * save area per workgroup): the same energies bit for bit, slower than
 * the default one workgroup per item. */
int libpt_amd_plan_geometry(const libpt_amd_plan *plan, int *out3);

/* enqueue ijk items [ijk_begin, ijk_end) of spin case sc on the plan's
 * stream; 0 on success */
int libpt_amd_plan_run(libpt_amd_plan *plan, int sc, int64_t ijk_begin,
    int64_t ijk_end);
int libpt_amd_plan_sync(libpt_amd_plan *plan);
/* energy of the call from the items run so far (all items: the full
 * energy); synchronises */
double libpt_amd_plan_energy(libpt_amd_plan *plan);
/* per-ijk energies of spin case sc (without the factor), n_ijk doubles */
int libpt_amd_plan_ijk_energies(libpt_amd_plan *plan, int sc, double *out);
/* the ijk triples of spin case sc, 3 * n_ijk ints, in work-list order */
int libpt_amd_plan_ijk_list(const libpt_amd_plan *plan, int sc, int32_t *out);

/* Tile geometry and order of the work list.  Tiles (A, B, C) are visited
 * in blocks of ga x gb x gc so that workgroups running together share
 * operand slices in L2; set_block applies one blocking.  set_geometry
 * switches the tile kernel between 4 waves (two workgroups per CU;
 * 32 x 32 x 16 tiles in FP64, 32 x 32 x 32 in the other forms) and 8
 * waves (one workgroup per CU; 32 x 32 x 32 tiles), rebuilding the plan's tile lists (not for the plans of a
 * batch).  tune times both geometries (only the plan's own when
 * LIBPT_AMD_WAVES is set) and the candidate blockings on the first
 * `sample_ijk` ijk triples of each spin case (0: 8), applies the fastest
 * and records it for plans of the same (form, call, sizes), also in the
 * file named by LIBPT_AMD_TUNE_CACHE when set; out4 (may be NULL)
 * receives (ga, gb, gc, waves).  LIBPT_AMD_TILE_BLOCK="ga,gb,gc" forces a
 * blocking for new plans.  0 on success. */
int libpt_amd_plan_set_block(libpt_amd_plan *plan, int ga, int gb, int gc);
int libpt_amd_plan_set_geometry(libpt_amd_plan *plan, int waves);
int libpt_amd_plan_tune(libpt_amd_plan *plan, int sample_ijk, int *out4);
/* records a geometry (waves 4 or 8; 0: the default) and blocking for new
 * plans of (call, form, sizes), as tune does after timing, e.g. from a
 * tuning table kept by the caller; 0 on success */
int libpt_amd_tune_record(int call, int form, size_t oa, size_t va,
    size_t ob, size_t vb, int ga, int gb, int gc, int waves);

/* GPU-timed window on the plan's stream: begin/end bracket enqueued work */
int libpt_amd_plan_timer_begin(libpt_amd_plan *plan);
double libpt_amd_plan_timer_end_ms(libpt_amd_plan *plan);

/* ---- batches of independent systems ---- */
typedef struct libpt_amd_batch libpt_amd_batch;

/* n independent systems of one call kind, in one arithmetic form.
 * dims[4 s .. 4 s + 3] = (oa, va, ob, vb) of system s (restricted calls
 * read oa, va); dev_arrays[NA s + t] = its t-th pt.h array in device memory
 * (NA = 7 for rpt/upt, 10 for rft/uft).  Stages every system's operands
 * and records one graph: the tile launches of every (system, spin case)
 * in one chain, largest items first, then one reduction per spin case.
 * NULL on error. */
libpt_amd_batch *libpt_amd_batch_create(int call, int form, int n,
    const int64_t *dims, const void *const *dev_arrays);
void libpt_amd_batch_destroy(libpt_amd_batch *batch);
/* enqueue one launch of the batch's graph (asynchronous); 0 on success */
int libpt_amd_batch_run(libpt_amd_batch *batch);
/* energies of the n systems from the last run (synchronises); each equals
 * libpt_amd_plan_energy of that system alone, bit for bit */
int libpt_amd_batch_energies(libpt_amd_batch *batch, double *energies);
/* GPU time of the last run in ms (synchronises) */
double libpt_amd_batch_last_ms(libpt_amd_batch *batch);
/* reference-formulation flops of one run, and its work items */
double libpt_amd_batch_flop(const libpt_amd_batch *batch);
int64_t libpt_amd_batch_items(const libpt_amd_batch *batch);
/* tile order of every system (as libpt_amd_plan_tune): systems whose sizes
 * have a recorded blocking take it, the others are tuned on the first
 * `sample_ijk` ijk triples of each spin case (0: 8) and recorded, so each
 * distinct size is timed once; later runs of the batch use the new orders.
 * Returns the number of systems timed, -1 on error.  The tuning runs
 * overwrite the batch's energies until its next run. */
int libpt_amd_batch_tune(libpt_amd_batch *batch, int sample_ijk);

/* ---- device memory and synthetic inputs ---- */
void *libpt_amd_dev_alloc(size_t bytes);
void libpt_amd_dev_free(void *ptr);
int libpt_amd_memcpy(void *dst, const void *src, size_t bytes, int kind);

/* Fill one pt.h input array with the deterministic generator (see
 * libpt_amd/csrc/synth.h).  spec points to a struct pt_synth_spec;
 * tensor is a pt_synth_tensor id; on_device selects device or host
 * memory; fp32 writes float instead of double. */
int libpt_amd_synth(const void *spec, int tensor, void *dst, int on_device,
    int fp32);
int64_t libpt_amd_synth_size(const void *spec, int tensor);
/* elements [begin, begin + count) of the same tensor into host memory
 * (dst[0] is element begin), e.g. one occupied slab of a large input */
int libpt_amd_synth_range(const void *spec, int tensor, void *dst,
    int64_t begin, int64_t count, int fp32);

/* copy n doubles device-to-device with a plain kernel (reference point
 * for hardware traffic counters) */
int libpt_amd_copy_kernel(void *dst, const void *src, int64_t n);

/* message of the calling thread's last failed call of this API */
const char *libpt_amd_last_error(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* LIBPT_AMD_H */
