/*
 * cuslines_hip.h -- C ABI of libcuslines_hip.so, the gfx950 tracking library.
 *
 * The reference (dipy/GPUStreamlines, package ``cuslines``) reaches its
 * tracking kernels from Python through cuda-python: one NVRTC-compiled module
 * per direction getter, two kernel launches per seed chunk, device buffers
 * owned by ``GPUTracker``.  This header is the plain C replacement of that
 * boundary.  Every entry point takes plain pointers and sizes, returns an int
 * status (0 = success) and leaves a message for cuslines_last_error().
 *
 * Threading: one host thread per tracker handle; each handle owns one HIP
 * stream (default priority); calls are synchronous unless documented as
 * "_async" (those enqueue on the handle's stream and return).
 */
#ifndef CUSLINES_HIP_H
#define CUSLINES_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUSLINES_ABI_VERSION 1

/* Direction getter / model kinds; values of the reference's ModelType enum
 * (cuslines/cuda_c/globals.h:82-87, cuslines/cuda_python/cutils.py:14-19). */
enum {
    CUSLINES_OPDT = 0,
    CUSLINES_CSA = 1,
    CUSLINES_PROB = 2,
    CUSLINES_PTT = 3
};

/* Tissue classes returned by the stop check (cuslines/cuda_c/globals.h:89). */
enum {
    CUSLINES_OUTSIDEIMAGE = 0,
    CUSLINES_INVALIDPOINT = 1,
    CUSLINES_TRACKPOINT = 2,
    CUSLINES_ENDPOINT = 3
};

/* Status codes. */
enum {
    CUSLINES_OK = 0,
    CUSLINES_EINVAL = 1,   /* bad argument (shape, range, null pointer)     */
    CUSLINES_EHIP = 2,     /* HIP runtime error                              */
    CUSLINES_ESTATE = 3,   /* call out of order (e.g. phase 2 before data)   */
    CUSLINES_ENOMEM = 4    /* device allocation failed                       */
};

/*
 * Tracking parameters.  Replaces the NVRTC macro set built in
 * cuslines/cuda_python/cu_direction_getters.py:70-110 (common macros) and
 * :456-468 (PTT macros), plus GPUTracker's constructor arguments
 * (cuslines/cuda_python/cu_tractography.py:37-56).  Float fields hold the
 * float32 values the reference rounds its arguments to.
 */
typedef struct cuslines_params {
    int32_t dimx, dimy, dimz, dimt;  /* data volume [X][Y][Z][T] (C order)      */
    int32_t model_type;              /* CUSLINES_OPDT/CSA/PROB/PTT              */
    int32_t nverts;                  /* sphere vertices (SAMPLM_NR)             */
    int32_t nedges;                  /* sphere edges (NUM_EDGES)                */
    int32_t sphere_symm;             /* SPHERE_SYMM                              */
    int32_t max_sline_len;           /* MAX_SLINE_LEN (501)                      */
    int32_t delta_nr;                /* bootstrap: SH coefficients               */
    uint64_t rng_seed;               /* RNG_SEED                                 */
    float step_size;                 /* STEP_SIZE, voxels                        */
    float max_angle;                 /* MAX_ANGLE, radians                       */
    float tc_threshold;              /* TC_THRESHOLD (stop-map threshold)        */
    float relative_peak_thresh;      /* RELATIVE_PEAK_THRESH                     */
    float min_separation_angle;      /* MIN_SEPARATION_ANGLE, radians            */
    float min_signal;                /* bootstrap min_signal                     */
    /* PTT only */
    int32_t step_frac;               /* STEP_FRAC                                */
    int32_t probe_quality;           /* PROBE_QUALITY (probe steps)              */
    int32_t lut_res;                 /* nearest-vertex LUT resolution (128)      */
    int32_t tile_log2_x;             /* LOG2_X of the (nx,ny,nz) vertex tiling   */
    int32_t tile_log2_y;             /* LOG2_Y                                   */
    int32_t tile_log2_z;             /* log2 nz                                  */
    float probe_step_size;           /* PROBE_STEP_SIZE                          */
    float max_curvature;             /* MAX_CURVATURE                            */
} cuslines_params;

/*
 * Learned PTT proposal (no counterpart in the reference).  At every
 * non-initial PTT sub-step the reference measures the probe support of the 24
 * disc vertices (8 probe taps each) to build its face-CDF proposal
 * (ptt.cu:232-300).  The learned proposal predicts those 24 supports with a
 * BF16 MLP 40 -> 128 -> 128 -> 24 (ReLU) from cheap local inputs, then samples
 * and accepts candidates with the reference's own face CDF and acceptance
 * rule (support >= 0.4 over 8 probes, lowest lane wins); a round without a
 * winner falls back to the plain getter.
 * Inputs: [0..23] for each disc vertex, the sum of two ODF taps along its
 * curve at arc lengths 4 and 8 probe steps; [24] the ODF at the point along
 * the tangent; [25..27] zero; [28..30] incoming direction; [31..39] frame.
 * Parameter layout (row-major, CUSLINES_MLP_NPARAM values): W1[128][40],
 * b1[128], W2[128][128], b2[128], W3[24][128], b3[24].
 * A training pair is CUSLINES_PAIR_FLOATS floats: the 40 inputs and the 24
 * measured vertex supports.
 */
#define CUSLINES_MLP_IN 40
#define CUSLINES_MLP_HIDDEN 128
#define CUSLINES_MLP_OUT 24
#define CUSLINES_MLP_NPARAM 24856
#define CUSLINES_PAIR_FLOATS 64

typedef struct cuslines_tracker cuslines_tracker;

/* ABI version of the loaded library (CUSLINES_ABI_VERSION). */
int cuslines_abi_version(void);

/* Copy the calling thread's last error message into buf (NUL-terminated).
 * Mirrors the message half of checkCudaErrors, cuslines/cuda_python/cutils.py:56-75. */
int cuslines_last_error(char *buf, size_t len);

/* Number of visible HIP devices (cuslines/cuda_python/cu_tractography.py:161-165). */
int cuslines_device_count(int *count);

/* Create a tracker bound to one device.  Replaces GPUTracker.__init__'s
 * context/stream setup and compile_program (cu_tractography.py:110-183,
 * cu_direction_getters.py:53-144): the kernels are precompiled for gfx950 and
 * specialised by the runtime parameters above. */
int cuslines_tracker_create(const cuslines_params *params, int device, cuslines_tracker **out);

/* Free every device buffer and the stream (cu_tractography.py:260-300). */
int cuslines_tracker_destroy(cuslines_tracker *t);

/* A second handle on the same device that shares t's device data (volume,
 * stop map, sphere, getter matrices, PTT LUT and blocks, proposal network)
 * and has its own stream and queue counters, so consecutive seed chunks can
 * run concurrently: a chunk's phase 2 fills the GPU while the previous
 * chunk's longest streamlines finish.  Replaces nothing in the reference,
 * whose chunks run one after the other on one stream per GPU
 * (cuslines/cuda_python/cu_propagate_seeds.py:206-233).  Configure t before
 * forking: calls that replace device data are refused on a fork and on a
 * tracker with live forks (CUSLINES_ESTATE); destroy forks before t.  The
 * fork copies t's launch and kernel settings and has no learned-proposal
 * statistics or pair recording of its own. */
int cuslines_tracker_fork(cuslines_tracker *t, cuslines_tracker **out);

/* Upload the data volume: DWI for bootstrap, ODF amplitudes for prob, the
 * prepared ODF (clip >= 0, zero outside the stop mask, max-normalised) for
 * PTT.  Host float32 [dimx][dimy][dimz][dimt] (cu_tractography.py:204-221).
 * PTT data with a negative value is accepted; rejection rounds then measure
 * every candidate's full support (the results are the same). */
int cuslines_set_data(cuslines_tracker *t, const float *dataf);

/* Upload the stop map, host float32 [dimx][dimy][dimz]
 * (cu_tractography.py:237-241: linear-filtered texture, border -1). */
int cuslines_set_stop_map(cuslines_tracker *t, const float *stop_map);

/* Upload sphere vertices float32 [nverts][3] and edges int32 [nedges][2]
 * (cu_tractography.py:223-255). */
int cuslines_set_sphere(cuslines_tracker *t, const float *vertices, const int32_t *edges);

/* Bootstrap matrices (BootDirectionGetter.allocate_on_gpu,
 * cu_direction_getters.py:217-280): H, R float32 [hr][hr] with hr the number
 * of non-b0 volumes; delta_b, delta_q float32 [delta_nr][hr];
 * sampling_matrix float32 [nverts][delta_nr]; b0s_mask int32 [dimt]. */
int cuslines_set_boot_matrices(cuslines_tracker *t, const float *H, const float *R,
                               const float *delta_b, const float *delta_q,
                               const float *sampling_matrix, const int32_t *b0s_mask);

/* PTT nearest-vertex lookup table, int32 [lut_res]^3 indexed [ix][iy][iz]
 * over linspace(-1, 1, lut_res) per axis (PttDirectionGetter.prepare_data,
 * cu_direction_getters.py:601-623, uploaded as a point-sampled texture). */
int cuslines_set_ptt_lut(cuslines_tracker *t, const int32_t *lut);

/* PTT data preparation on the device, in place of cuslines_set_data plus
 * cuslines_set_ptt_lut with host-prepared inputs: PttDirectionGetter.prepare_data
 * (cu_direction_getters.py:531-625) clips the raw ODF `odf` (host float32
 * [dimx][dimy][dimz][dimt]) at 0, zeroes voxels whose `stop_map` value is
 * below `stop_threshold`, divides each voxel by its maximum when positive,
 * and builds the nearest-vertex LUT of the uploaded sphere (antipodes too
 * when `sphere_symm`) over linspace(-1, 1, lut_res)^3.  The sphere must be
 * uploaded first; the tiling comes from the tracker's parameters.  `lut_out`
 * (optional, int32 [lut_res]^3) receives the LUT.  Same results as the host
 * recipe. */
int cuslines_ptt_prepare(cuslines_tracker *t, const float *odf, const float *stop_map, float stop_threshold,
                         int sphere_symm, int32_t *lut_out);

/* Phase-2 schedules. */
enum {
    CUSLINES_SCHED_QUEUE = 0,  /* persistent grid, slots pull seeds from a counter */
    CUSLINES_SCHED_STATIC = 1  /* one seed per 32-lane slot, one launch-wide grid  */
};

/* Launch configuration of phase 2 (chosen by gpustreamlines_amd.autotune):
 * schedule, threads per block (0 = getter default, else 64..256 in steps of
 * 64; every 32 threads are one seed slot) and resident blocks per CU of the
 * persistent grid (0 = the occupancy maximum).  The reference's launch shape
 * is fixed at 2 seeds per 64-thread block (cu_propagate_seeds.py:45-54). */
int cuslines_set_launch(cuslines_tracker *t, int schedule, int threads, int blocks_per_cu);

/* Bootstrap (OPDT/CSA) phase-2 kernels. */
enum {
    CUSLINES_BOOT_EXACT = 0,  /* one seed per 32 lanes, the reference's lane-strided
                                 dot products: bit-identical to the CPU oracle    */
    CUSLINES_BOOT_MFMA = 1,   /* 16 seeds per workgroup in lockstep, the refit and
                                 ODF projection as fp32 MFMA products batched over
                                 the 16 seeds: equal within float rounding         */
    CUSLINES_BOOT_FULL_PRODUCTS = 0x10000  /* flag: MFMA mode with the full
                                 [H; R] product even where the low-rank form holds */
};

/* Select the bootstrap phase-2 kernel (default CUSLINES_BOOT_EXACT).  Replaces
 * genStreamlinesMergeBoot_k's per-seed ndotp_d products
 * (cuslines/cuda_c/boot.cu:195-235, 466-534) with batched matrix-core products
 * in CUSLINES_BOOT_MFMA mode; the algorithm, random streams and outputs'
 * layout are the same.  MFMA mode needs hr <= 320 DWI volumes, delta_nr <= 128
 * and nverts <= 768 (phase 2 returns CUSLINES_EINVAL otherwise) and the
 * queue schedule (CUSLINES_SCHED_STATIC keeps the exact kernel).  Bits 8..15
 * of `mode` choose the MFMA kernel's occupancy, 2 or 4 (default) waves per
 * SIMD (the autotuner's knob).
 * When H is a projection H = U U^T of rank r <= 64 and R its leveraged centred
 * residual matrix (boot_utils.py:86-94: R = L (I - H) - column means, L =
 * diag(1 / sqrt(1 - H_ii))), as cuslines_set_boot_matrices checks to 1e-5, the
 * MFMA kernel computes H.d = U (U^T d) and R.d = v - mean(v), v = L (d - H.d),
 * instead of the full [H; R] product (fewer products, another rounding);
 * CUSLINES_BOOT_FULL_PRODUCTS in `mode` keeps the full product. */
int cuslines_set_boot_mode(cuslines_tracker *t, int mode);

/* The rank r of H that the MFMA bootstrap kernel's low-rank form uses, or 0
 * when H and R do not have that form or CUSLINES_BOOT_FULL_PRODUCTS is set.
 * No counterpart in the reference. */
int cuslines_boot_lowrank(cuslines_tracker *t, int32_t *rank);

/* The same check on host arrays, without a tracker or a GPU: *rank = r when
 * H (hr x hr, row-major) is U U^T of rank r <= 64 < hr and R its leveraged
 * centred residual matrix, to 1e-5; else 0.  No counterpart in the
 * reference. */
int cuslines_hat_lowrank(const float *H, const float *R, int32_t hr, int32_t *rank);

/* Work counters of the MFMA bootstrap kernel since the last reset: out[0]
 * getter attempts of all slots (each one refit and ODF projection, i.e. one
 * column of the batched products), out[1] lockstep rounds of all workgroups
 * (each round runs the products for 16 columns); reset != 0 zeroes them.
 * attempts / (16 rounds) is the occupancy of the batched products. */
int cuslines_boot_counters(cuslines_tracker *t, int64_t *out2, int reset);

/* PTT phase 2 with `lanes` (8, 16 or 32) lanes per seed, 64/lanes seeds per
 * wavefront.  The reference's 32 per-lane random streams per seed are kept as
 * counter-based virtual lanes, so the output does not depend on this choice.
 * 0 (default) selects the one-seed-per-32-lane kernel shared with the other
 * getters (which always use 32 lanes per seed). */
int cuslines_set_lanes_per_seed(cuslines_tracker *t, int lanes);

/* Phase 2 on the work queue (the probabilistic getter, the exact bootstrap
 * kernel and plain PTT with lanes per seed 0): flat = 1 (default) runs each
 * 32-lane seed slot as one loop of steps (PTT: sub-steps) in which a slot
 * whose streamline pass ends starts its next pass, direction or seed at
 * once, so the two slots of a wavefront keep stepping together; flat = 0
 * runs the reference's nesting (genStreamlinesMerge{Boot,Prob}_k: per seed,
 * per direction, per pass, per step), where a slot whose loop ends waits
 * for the other slot's.  Same output either way.  No counterpart in the
 * reference. */
int cuslines_set_stepping(cuslines_tracker *t, int flat);

/* PTT probe taps: 1 (default) reads each tap's 8 texels as one 32-byte
 * corner block, an expanded copy of the ODF built once both
 * cuslines_set_data and cuslines_set_ptt_lut have been called, when HBM
 * allows ((X+1)(Y+1)(Z+1) x nverts x 32 B, cell-major); 0 gathers the 8
 * texels from the vertex-major ODF.  Results are identical either way. */
int cuslines_set_ptt_layout(cuslines_tracker *t, int corner_blocks);

/* Learned proposal for PTT phase 2: host bf16 parameters (uint16 bit
 * patterns, CUSLINES_MLP_NPARAM of them, layout above); NULL returns to the
 * plain getter.  A streamline whose learned rounds fail more than one time in
 * four (after 16 sub-steps) continues with the plain getter.  No counterpart
 * in the reference: it trades the 24 x 8 vertex probe taps of each sub-step
 * for 24 x 2 taps and the network. */
int cuslines_set_proposal(cuslines_tracker *t, const uint16_t *mlp_bf16);

/* Record training pairs during later plain PTT phase-2 calls: each
 * non-initial sub-step is kept with probability `rate` (a hash of seed, pass
 * and sub-step), appended at atomicAdd(count_d) into pairs_d
 * [capacity][CUSLINES_PAIR_FLOATS] while there is room (count_d keeps
 * counting past capacity).  pairs_d == NULL stops recording.  No counterpart
 * in the reference. */
int cuslines_set_recording(cuslines_tracker *t, float *pairs_d, int32_t capacity, int32_t *count_d,
                           float rate);

/* Proposal counters since the last reset: out[0] learned sub-steps accepted
 * in their learned round, out[1] learned rounds that fell back to the plain
 * getter, out[2] plain sub-steps of streamlines that left the learned
 * proposal.  reset != 0 zeroes them after reading. */
int cuslines_proposal_stats(cuslines_tracker *t, int64_t *out3, int reset);

/* The same three counters and out4[3]: face draws of the learned rounds
 * whose exact face value exceeded the envelope (PROPOSAL_M x the proposal's
 * face value), i.e. where the exact correction could not hold; 0 means the
 * learned sampler drew from the reference's distribution. */
int cuslines_proposal_counters(cuslines_tracker *t, int64_t *out4, int reset);

/* Instrumentation of PTT phase 2 (no counterpart in the reference): while
 * enabled, phase 2 runs an instrumented copy of the plain getter (same
 * results) that counts every ODF tap it reads (one LUT entry + one 8-texel
 * block each) and every sub-step.  cuslines_tap_count returns out[0] taps and
 * out[1] sub-steps since the last reset (reset != 0 zeroes them); it is
 * refused while a learned proposal is set, whose statistics share the
 * counters.  Currently switched off: enable != 0 returns CUSLINES_ESTATE
 * (the instrumented kernel stopped responding on its first GPU run and the
 * cause is not yet found); enable == 0 and cuslines_tap_count still work. */
int cuslines_set_tap_count(cuslines_tracker *t, int enable);
int cuslines_tap_count(cuslines_tracker *t, int64_t *out2, int reset);

/* Training of the proposal network on device buffers (no counterpart in the
 * reference).  params_d: float32 master parameters [CUSLINES_MLP_NPARAM];
 * the forward pass rounds them to bf16 as inference does.
 * cuslines_mlp_loss_grad_async: mean over n pairs of the squared error of
 * the 24 predicted supports, averaged over the 24, into loss_d[0], and its
 * gradient into grad_d [CUSLINES_MLP_NPARAM]; work_d is a caller buffer of
 * cuslines_mlp_workspace_bytes(n) bytes; n <= 8,388,480 (65535 chunks of
 * 128 pairs), larger n is rejected with CUSLINES_EINVAL.
 * cuslines_mlp_adam_async: one Adam step (Kingma and Ba, 2015) with step
 * count `step` (>= 1); m_d and v_d start at zero. */
int cuslines_mlp_workspace_bytes(int32_t n, size_t *bytes);
int cuslines_mlp_loss_grad_async(cuslines_tracker *t, const float *params_d, const float *pairs_d, int32_t n,
                                 float *loss_d, float *grad_d, void *work_d);
int cuslines_mlp_adam_async(cuslines_tracker *t, float *params_d, const float *grad_d, float *m_d, float *v_d,
                            int32_t step, float lr, float beta1, float beta2, float eps);

/* Device memory helpers for callers that keep seeds and outputs resident:
 * the reference's cudaMalloc / cudaMemcpy / cudaFree / cudaStreamSynchronize
 * calls on its per-GPU streams (cuslines/cuda_python/cu_propagate_seeds.py:56-201,220-231).
 * The copies are synchronous on the handle's stream. */
int cuslines_device_alloc(cuslines_tracker *t, size_t bytes, void **dptr);
int cuslines_device_free(cuslines_tracker *t, void *dptr);
int cuslines_memcpy_htod(cuslines_tracker *t, void *dst, const void *src, size_t bytes);
int cuslines_memcpy_dtoh(cuslines_tracker *t, void *dst, const void *src, size_t bytes);
int cuslines_synchronize(cuslines_tracker *t);
/* The handle's HIP stream (hipStream_t), for event timing by the caller. */
int cuslines_get_stream(cuslines_tracker *t, void **stream);

/*
 * Phase 1 on device buffers (getNumStreamlines{Boot,Prob,Ptt}_k,
 * cuslines/cuda_c/boot.cu:610-718, generate_streamlines_cuda.cu:423-462,
 * ptt_init.cu:1-63): for every seed, the initial peak directions.
 *   seeds_d    float32 [nseeds][3] voxel coordinates
 *   counts_d   int32 [nseeds + 1]; counts_d[i] = number of directions of seed i
 *   dirs_d     int32 [nseeds][nverts]; sphere-vertex index of each direction
 * rng_chunk: seeds are numbered i % rng_chunk for their random streams
 * (the reference's streams are chunk-relative); 0 means nseeds.
 */
int cuslines_phase1_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds,
                          int32_t rng_chunk, int32_t *counts_d, int32_t *dirs_d);

/* Exclusive scan of counts_d[0..nseeds) into offsets_d[0..nseeds] on device,
 * replacing the host loop of cuslines/cuda_python/cu_propagate_seeds.py:95-129.
 * The total is written to offsets_d[nseeds]. */
int cuslines_scan_async(cuslines_tracker *t, const int32_t *counts_d, int32_t nseeds,
                        int32_t *offsets_d);

/*
 * Phase 2 on device buffers (genStreamlinesMerge{Boot,Prob}_k,
 * cuslines/cuda_c/boot.cu:826-980, generate_streamlines_cuda.cu:464-622):
 * for every seed direction, track backward, reverse, track forward.
 *   offsets_d  int32 [nseeds + 1] exclusive offsets (from cuslines_scan_async)
 *   dirs_d     as written by phase 1
 *   sline_seed_d int32 [nslines]   seed index of each streamline
 *   sline_len_d  int32 [nslines]   points per streamline
 *   sline_d      float32 [nslines][2*max_sline_len][3] padded point slots
 */
int cuslines_phase2_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds,
                          int32_t rng_chunk, const int32_t *offsets_d,
                          const int32_t *dirs_d, int32_t *sline_seed_d,
                          int32_t *sline_len_d, float *sline_d);

/*
 * Phase 2 without allocation: cuslines_phase2_async with the same arguments,
 * kernels and results, except that it calls no hipMalloc / hipMallocAsync /
 * hipFree of its own.  No counterpart in the reference; it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid allocation).  The spatial seed order that the
 * persistent queue reads (order, keys, bucket counts, scan block sums) lives
 * in a static device array of the library (about 24 MiB, one per device,
 * held until the process ends) instead of stream-ordered allocations.
 * nseeds above CUSLINES_NOALLOC_MAX_SEEDS is refused with CUSLINES_EINVAL.
 * Every call on one device shares that array: calls that overlap on one
 * device (several trackers or streams) race and must be serialised by the
 * caller.
 */
#define CUSLINES_NOALLOC_MAX_SEEDS 1048576
int cuslines_phase2_noalloc_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds,
                                  int32_t rng_chunk, const int32_t *offsets_d,
                                  const int32_t *dirs_d, int32_t *sline_seed_d,
                                  int32_t *sline_len_d, float *sline_d);

/* Compact padded slots into one point array: points_d[sum(len)][3], with
 * point_off_d int64 [nslines + 1] the exclusive scan of sline_len_d. */
int cuslines_compact_async(cuslines_tracker *t, const int32_t *sline_len_d, int32_t nslines,
                           const float *sline_d, int64_t *point_off_d, float *points_d);

/*
 * One chunk of the reference's protocol without its host round trip
 * (cuslines/cuda_python/cu_propagate_seeds.py:95-201 reads the streamline
 * count back between phase 1 and phase 2 to size the output slots): phase 1,
 * the scan, phase 2 and the compaction are enqueued on the handle's stream
 * against a fixed slot capacity.  The call reads nothing back and allocates
 * nothing (the workspace is the caller's), so it can be captured in a HIP
 * graph (cuslines_chunk_graph_create).  No counterpart in the reference.
 *   seeds_d, counts_d, dirs_d, offsets_d  as cuslines_phase1_async and
 *                cuslines_scan_async (counts_d [nseeds + 1], offsets_d [nseeds + 1])
 *   slot_cap     slots of sline_seed_d, sline_len_d and sline_d (>= 1)
 *   point_off_d  int64 [slot_cap + 1]: exclusive scan of the slot lengths
 *   points_d     float32 [slot_cap * (2 * max_sline_len - 1)][3] (the most the
 *                slots can hold); the compacted points of the tracked streamlines
 *   totals_d     int64 [3]: streamlines of the chunk (offsets_d[nseeds]), slots
 *                tracked, points compacted
 *   ws_d         cuslines_chunk_workspace_bytes(nseeds, slot_cap) bytes,
 *                256-byte aligned (cuslines_device_alloc memory is)
 * Phase 2 tracks the seeds whose streamlines all lie below slot_cap -- a
 * prefix of the seeds, as slots follow seed order -- with the results
 * cuslines_phase2_async gives them; every other slot reads length 0.  When
 * totals_d[0] > totals_d[1] the chunk did not fit: run it again with
 * slot_cap >= totals_d[0].
 */
int cuslines_chunk_workspace_bytes(cuslines_tracker *t, int32_t nseeds, int32_t slot_cap, size_t *bytes);
int cuslines_chunk_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                         int32_t *counts_d, int32_t *dirs_d, int32_t *offsets_d, int32_t slot_cap,
                         int32_t *sline_seed_d, int32_t *sline_len_d, float *sline_d, int64_t *point_off_d,
                         float *points_d, int64_t *totals_d, void *ws_d, size_t ws_bytes);

/*
 * HIP graph of one cuslines_chunk_async call with fixed arguments, captured
 * on the handle's stream; it replaces the per-chunk kernel launches of
 * cuslines/cuda_python/cu_propagate_seeds.py:206-233.  The buffers' contents
 * may change between launches (copy the next chunk's seeds into seeds_d,
 * then launch); the launches run as the handle was set up at capture (getter
 * data, launch shape, proposal).  cuslines_graph_launch enqueues the graph on
 * the handle's stream.  While a graph lives, its tracker (and a fork's
 * parent) refuses to be destroyed or to take new data (CUSLINES_ESTATE).
 * No counterpart in the reference.
 */
typedef struct cuslines_graph cuslines_graph;
int cuslines_chunk_graph_create(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                                int32_t *counts_d, int32_t *dirs_d, int32_t *offsets_d, int32_t slot_cap,
                                int32_t *sline_seed_d, int32_t *sline_len_d, float *sline_d, int64_t *point_off_d,
                                float *points_d, int64_t *totals_d, void *ws_d, size_t ws_bytes,
                                cuslines_graph **out);
int cuslines_graph_launch(cuslines_graph *g);
int cuslines_graph_destroy(cuslines_graph *g);

#ifdef __cplusplus
}
#endif

#endif /* CUSLINES_HIP_H */
