/* rsim: the MI355X (gfx950) batched RocketSim V3 decision stepper — C ABI of librsim.so.
 *
 * Drop-in for the reference crate's device API (crates/rocketsim-cuda/src/device.rs:15-172, SimDevice and its buffers).
 * Conventions of every entry point:
 *   - returns an int status: RSIM_OK (0) or a negative RSIM_ERR_*; never throws, never aborts on misuse (the reference
 *     panics with assert! where this returns RSIM_ERR_INVALID / RSIM_ERR_SHORT);
 *   - handles are opaque and owned by the caller, freed explicitly (rsim_*_free, rsim_close);
 *   - host memory is caller-owned; `ArenaState`, `StaticData` and `Controls` are passed as `void*` to buffers in the
 *     reference's word format (cuda/sim/arena_state.cuh:38-342: ArenaState 11,136 B, StaticData 1,271,580 B, Controls
 *     32 B = throttle, steer, pitch, yaw, roll as f32, then jump, boost, handbrake as u32 0/1);
 *   - all device work of one rsim_device goes to its one stream, in call order: rsim_sim8 and rsim_run are asynchronous,
 *     everything that returns host data synchronises first. A device handle is not thread-safe;
 *   - while the stream records a graph (rsim_capture_begin .. rsim_capture_end) every call that allocates, frees, copies
 *     to or from the host, times or synchronises (rsim_alloc_*, rsim_upload_*, rsim_write_*, rsim_read_words,
 *     rsim_download_*, rsim_*_free, rsim_sync, rsim_timer_*, the parity kernels, rsim_close) returns RSIM_ERR_INVALID
 *     and changes nothing: allocate every buffer a captured sequence uses before the capture.
 * rsim_last_error() describes the last failure of the calling thread. */
#ifndef RSIM_H
#define RSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------------------------------- */
#define RSIM_OK 0
#define RSIM_ERR_INVALID (-1)   /* bad argument: null pointer, wrong size, block outside 1..RSIM_MAX_BLOCK ... */
#define RSIM_ERR_SHORT (-2)     /* a buffer is shorter than the call needs (device.rs:158, :163) */
#define RSIM_ERR_HIP (-3)       /* a HIP runtime call failed (DriverError in the reference) */
#define RSIM_ERR_NO_DEVICE (-4) /* no such GPU, or no GPU at all */

/* PhaseStatus of the per-arena status words (cuda/sim/phase.cuh:18). */
#define RSIM_PHASE_OK 0
#define RSIM_PHASE_UNIMPLEMENTED (-1)
#define RSIM_PHASE_BAD_INDEX (-2)

/* ---- the decision and its result word (device.rs:15-28; cuda/sim/arena_step.cuh:16-19) ----------------------------- */
#define RSIM_TICKS_PER_DECISION 8
#define RSIM_SIM8_SCORED_MASK 0xFFu         /* the tick 1..8 after which the ball was first scored, 0 for none */
#define RSIM_SIM8_UNIMPLEMENTED 0x80000000u /* a tick up to then reached a shape pair V3's port does not implement */
#define RSIM_SIM8_OVERFLOW 0x40000000u      /* an overflow counter of the state written back is nonzero */
#define RSIM_SIM8_BAD_VARIANT 0x20000000u   /* (rsim_sim8_variants) the arena's world index is past the table: not stepped */
#define RSIM_SIM8_BLOCK 64                  /* work-group size in production: one wave64 per AoSoA-64 block */
#define RSIM_MAX_BLOCK 256                  /* largest work-group size the tick kernels accept */
#define RSIM_NUM_PHASES 10                  /* P0, P1, P2, P3a, P3b, P4, P5, P6, P7, P8 (phase.cuh:17-46) */
#define RSIM_AOSOA_WIDTH 64                 /* arenas per AoSoA block */

/* ---- layout queries (crates/rocketsim-cuda/csrc/phase.cpp:18-22 lf_*; src/aosoa.rs:1-38) --------------------------- */
uint64_t rsim_arena_words(void);  /* 2,784 u32 words per ArenaState */
uint64_t rsim_tick_words(void);   /* 2,567: the words that persist across a tick boundary; the rest is zero there */
uint64_t rsim_static_words(void); /* 317,895 u32 words per StaticData */
/* u32 words of an AoSoA-64 buffer of n arenas: ceil(n / 64) blocks of 2,784 x 64 words. */
uint64_t rsim_aosoa_words(uint32_t n);
/* Position of word w of arena a: (a / 64) * 2784 * 64 + w * 64 + a % 64. */
uint64_t rsim_aosoa_index(uint32_t a, uint32_t w);
/* Host-side pack / unpack between n ArenaStates and an AoSoA-64 buffer (aosoa.rs:17-38); pack zero-fills padding lanes. */
int rsim_aosoa_pack(const void* states, uint32_t n, uint32_t* aosoa);
int rsim_aosoa_unpack(const uint32_t* aosoa, uint32_t n, void* states);

/* ---- the device (device.rs:30-67) ------------------------------------------------------------------------------------ */
typedef struct rsim_device rsim_device;
typedef struct rsim_world rsim_world;   /* StaticData on the device (DeviceStatic, device.rs:47-50) */
typedef struct rsim_states rsim_states; /* n arenas in one AoSoA-64 device buffer (DeviceStates, device.rs:41-45) */
typedef struct rsim_words rsim_words;   /* a plain u32 device buffer (controls, results, traces: CudaSlice<u32>) */

/* SimDevice::new (device.rs:53-67): opens GPU `ordinal`, creates its stream (default priority) and timing events.
 * RSIM_ERR_HIP also if the library's stepping kernels declare static LDS (their LDS addressing assumes none). */
int rsim_open(int ordinal, rsim_device** dev);
/* Releases the device. Every world, state buffer, word buffer (views included) and graph made on it must be freed
 * first: while any is open this returns RSIM_ERR_INVALID and the device stays usable (the reference's Rust ownership
 * makes the order impossible to get wrong; a C caller gets the refusal instead of a freed device behind a live handle). */
int rsim_close(rsim_device* dev);
/* The device's HIP stream (a hipStream_t) and its ordinal, for callers that order their own work against it. */
void* rsim_stream(rsim_device* dev);

/* upload_static (device.rs:69-71): one StaticData (rsim_static_words() words). */
int rsim_upload_world(rsim_device* dev, const void* static_data, rsim_world** world);
/* No counterpart in the reference (one StaticData per launch, device.rs:155): a table of `count` StaticData back to back
 * (count <= RSIM_MAX_WORLDS), for batches that mix car pairings or mutators; rsim_sim8_variants picks one per arena. */
#define RSIM_MAX_WORLDS 1024
int rsim_upload_worlds(rsim_device* dev, const void* static_data, uint32_t count, rsim_world** world);
uint32_t rsim_world_count(const rsim_world* world);
/* The device address of the world table's first StaticData (for user kernels on include/rsim_device.h). */
const void* rsim_world_device_ptr(const rsim_world* world);
int rsim_world_free(rsim_world* world);

/* upload_states / alloc_states (device.rs:73-80): n ArenaStates packed to AoSoA-64 on the host and uploaded; alloc zeroes. */
int rsim_upload_states(rsim_device* dev, const void* states, uint32_t n, rsim_states** out);
int rsim_alloc_states(rsim_device* dev, uint32_t n, rsim_states** out);
/* Overwrites the n states of an existing buffer (same n) without reallocating. */
int rsim_write_states(rsim_device* dev, rsim_states* dst, const void* states);
/* Device-to-device copy of a whole state buffer (same n): the restore between timed runs (tests/speed.rs:65-75). */
int rsim_copy_states(rsim_device* dev, rsim_states* dst, const rsim_states* src);
int rsim_states_free(rsim_states* states);
uint32_t rsim_states_count(const rsim_states* states);
void* rsim_states_device_ptr(rsim_states* states);
/* download_words / download_states (device.rs:82-88): synchronous. */
int rsim_download_words(rsim_device* dev, const rsim_states* states, uint32_t* aosoa_out);
int rsim_download_states(rsim_device* dev, const rsim_states* states, void* states_out);

/* Plain u32 device buffers. upload_controls (device.rs:143-149) is rsim_upload_controls: Controls[n][2] (16 words per arena). */
int rsim_alloc_words(rsim_device* dev, uint64_t words, rsim_words** out);
int rsim_upload_words(rsim_device* dev, const uint32_t* host, uint64_t words, rsim_words** out);
int rsim_write_words(rsim_device* dev, rsim_words* dst, const uint32_t* host, uint64_t words);
int rsim_read_words(rsim_device* dev, const rsim_words* src, uint32_t* host, uint64_t words);
int rsim_upload_controls(rsim_device* dev, const void* controls, uint32_t n, rsim_words** out);
int rsim_words_free(rsim_words* buf);
/* Words offset..offset+words-1 of `parent` as a buffer of their own, for any call that takes a words buffer (no copy;
 * `parent` must outlive it; freeing the view leaves the memory alone). No counterpart in the reference: it lets a
 * caller step through one device-resident array in slices, e.g. the cost model's training batches. */
int rsim_words_view(rsim_words* parent, uint64_t offset, uint64_t words, rsim_words** out);
uint64_t rsim_words_count(const rsim_words* buf);
void* rsim_words_device_ptr(rsim_words* buf);

/* sim8 (device.rs:151-168): one decision (8 ticks) of every arena of `states`, in place, asynchronously on the device's
 * stream. controls: >= n * 16 words (rsim_upload_controls). result: >= n words, result[a] the RSIM_SIM8_* word of arena a.
 * block: work-group size, 1..RSIM_MAX_BLOCK (RSIM_SIM8_BLOCK in production); results do not depend on it. trace: null,
 * or >= 8 * rsim_aosoa_words(n) words that receive the state after tick t in slab t - 1 (ticks after a goal are not
 * written). */
int rsim_sim8(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* controls, rsim_words* result,
              uint32_t block, rsim_words* trace);
/* sim8 over a world table: arena a steps in world variant[a] (variant: >= n words, or null for world 0 everywhere). An
 * index >= rsim_world_count(world) leaves the arena untouched and sets RSIM_SIM8_BAD_VARIANT in its result word. No
 * counterpart in the reference; with a one-world table and a null variant it is rsim_sim8. */
int rsim_sim8_variants(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                       const rsim_words* controls, rsim_words* result, uint32_t block, rsim_words* trace);
/* rsim_sim8_variants that also writes each lane's wall-clock ticks (100 MHz) for its decision into `cycles` (>= 2 * n
 * words: n u64). The lanes of a wave share its clock; launched with block 1 every arena has a wave of its own and the
 * count is that arena's own cost (the cost model's labels). No counterpart in the reference. */
int rsim_sim8_timed(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* result, uint32_t block, rsim_words* trace, rsim_words* cycles);

/* `decisions` consecutive decisions in one launch, asynchronous: the same results as that many rsim_sim8_variants calls
 * (decision d + 1 starts from decision d's written-back state), with each arena held on chip from the first decision to
 * the last instead of being copied out and in between. controls: decisions x n x 16 words (decision-major); results:
 * decisions x n words. variant: as in rsim_sim8_variants (null: world 0). No counterpart in the reference; for callers
 * whose controls are known ahead (replays, scripted opponents, benchmarks). */
int rsim_run(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant, const rsim_words* controls,
             rsim_words* results, uint32_t decisions, uint32_t block);

/* rsim_run with dynamic scheduling: `workgroups` persistent work-groups of `block` lanes (64, 128, 192 or 256: whole
 * waves; workgroups 0 = as many as the GPU holds at once, rsim_run_queue_capacity) whose waves each take the next AoSoA
 * block of 64 arenas from a queue and run all `decisions` decisions of it. The queue holds the ceil(n / 64) blocks in
 * the order `order` (a permutation of the block indices, e.g. rsim_cost_block_order's heaviest-first order) or in index
 * order (order null). Same results and states as rsim_run: arenas are independent and each is stepped by one lane.
 * Asynchronous; calls on one device are ordered by its stream (the queue counter is the device's). No counterpart in
 * the reference. */
int rsim_run_queued(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* results, uint32_t decisions, uint32_t block, const rsim_words* order,
                    uint32_t workgroups);
uint32_t rsim_run_queue_capacity(rsim_device* dev, uint32_t block); /* resident work-groups of that size (0: unknown) */
/* One decision as ceil(8 / ticks_per_launch) launches of ticks_per_launch ticks, the state written back to the AoSoA
 * buffer between launches (ticks_per_launch 8 is one launch, as rsim_sim8_variants). Same result words and states as
 * rsim_sim8_variants; a schedule the autotuner measures. Asynchronous. No counterpart in the reference (its sim8 is one
 * launch per decision; step_tick is its one-tick debug kernel, device.rs:96-141). */
int rsim_sim8_split(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* result, uint32_t block, uint32_t ticks_per_launch);

/* ---- the learned cost model and arena regrouping (no counterpart in the reference) ----------------------------------------
 * A 48 -> 64 -> 64 -> 2 MLP over per-arena features predicts (log of the arena's decision cost, contact logit). Inference
 * runs on bf16 MFMA with fp32 accumulation; training is the library's own forward/backward/Adam kernels. Parameters: a
 * u32 buffer of RSIM_COST_PARAMS fp32 words: W1[48][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]. All asynchronous. */
#define RSIM_COST_FEATURES 48
#define RSIM_COST_PARAMS 7426
#define RSIM_COST_ACT_WORDS 306 /* training scratch words per batch row */
int rsim_cost_features(rsim_device* dev, const rsim_states* states, rsim_words* features); /* >= n * 48 fp32 */
int rsim_cost_predict(rsim_device* dev, const rsim_words* params, const rsim_words* features, uint32_t n, rsim_words* out); /* out: n x 2 */
/* One batch: loss[i] per row (squared error on target + logistic loss on contact) and grad = d(mean loss)/d(params).
 * target, contact: batch fp32; scratch: batch * RSIM_COST_ACT_WORDS words. */
int rsim_cost_gradients(rsim_device* dev, const rsim_words* params, const rsim_words* features, const rsim_words* target,
                        const rsim_words* contact, uint32_t batch, rsim_words* scratch, rsim_words* loss, rsim_words* grad);
int rsim_cost_adam(rsim_device* dev, rsim_words* params, rsim_words* m, rsim_words* v, const rsim_words* grad, float lr, uint32_t step);
/* The order of n arenas by predicted cost, computed on the device with no host synchronisation: perm[i] = the current
 * slot that moves to slot i. pred: n x 2 fp32 from rsim_cost_predict (column 0 read). The predicted normalised log cost
 * is bucketed (RSIM_COST_ORDER_BUCKETS buckets of 1/64 below 3.0, above 3.0 in the first, NaN in the last) and the slots
 * are sorted by bucket, most expensive first, stably (ties keep slot order), so the result is deterministic.
 * RSIM_ORDER_SORTED: that order (cost-homogeneous waves of 64). RSIM_ORDER_BALANCED: the sorted ranks dealt round-robin
 * over the ceil(n / 64) waves (every wave gets one arena of each cost rank; the last, partial wave only in the first
 * n - 64 (waves - 1) rounds). work: rsim_cost_order_work_words(n) words of scratch. */
#define RSIM_COST_ORDER_BUCKETS 256
#define RSIM_ORDER_SORTED 0
#define RSIM_ORDER_BALANCED 1
uint64_t rsim_cost_order_work_words(uint32_t n);
int rsim_cost_order(rsim_device* dev, const rsim_words* pred, uint32_t n, int mode, rsim_words* work, rsim_words* perm);
/* The order of the ceil(n / 64) AoSoA blocks by their largest predicted arena cost (pred as for rsim_cost_order), most
 * expensive first, stable: the queue order of rsim_run_queued (longest work first). work: rsim_cost_block_order_work_words(n). */
uint64_t rsim_cost_block_order_work_words(uint32_t n);
int rsim_cost_block_order(rsim_device* dev, const rsim_words* pred, uint32_t n, rsim_words* work, rsim_words* order);
/* rsim_cost_order without the work buffer, for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops that forbid hipMalloc). No counterpart in the reference. The workspace is a fixed 1 MiB array inside
 * librsim.so, one per GPU, held until the process ends; n must be <= RSIM_COST_ORDER_NOALLOC_MAX (else RSIM_ERR_INVALID),
 * and the call allocates nothing. Same results as rsim_cost_order. Every call on one GPU shares that GPU's array: two
 * orders in flight at once on one GPU (on different streams, or from two rsim_device handles of the same GPU) race. */
#define RSIM_COST_ORDER_NOALLOC_MAX 262144u
int rsim_cost_order_noalloc(rsim_device* dev, const rsim_words* pred, uint32_t n, int mode, rsim_words* perm);
/* Uploads a host permutation of 0..n-1 (checked) for the regrouping calls. */
int rsim_set_permutation(rsim_device* dev, const uint32_t* perm, uint32_t n, rsim_words* dst);
/* dst arena i = src arena perm[i] for the dst->n arenas of dst (whole AoSoA-64 states; src and dst distinct; perm holds
 * at least dst->n indices; an index past src leaves that arena of dst as it was). With dst->n == src->n and a permutation
 * this is a reordering; with a smaller dst, a sample. */
int rsim_gather_states(rsim_device* dev, const rsim_states* src, const rsim_words* perm, rsim_states* dst);
/* Items of words_per_item words: gather (dst[i] = src[perm[i]]) or, with scatter != 0, dst[perm[i]] = src[i]. */
int rsim_permute_words(rsim_device* dev, const rsim_words* src, const rsim_words* perm, uint32_t n, uint32_t words_per_item,
                       int scatter, rsim_words* dst);

/* The debug and parity kernels (device.rs:90-141), synchronous. status_out: n ints (PhaseStatus per arena) or null. */
int rsim_identity(rsim_device* dev, const rsim_states* in, rsim_states* out);      /* copy-in, copy-out */
int rsim_identity_tick(rsim_device* dev, const rsim_states* in, rsim_states* out); /* tick-boundary copy: tail zeroed */
int rsim_phase(rsim_device* dev, int k, rsim_states* states, const rsim_world* world, int32_t* status_out);
int rsim_step_tick(rsim_device* dev, rsim_states* states, const rsim_world* world, int32_t* status_out);

/* HIP graphs over the device's stream: between rsim_capture_begin and rsim_capture_end the asynchronous entry points
 * (rsim_sim8*, rsim_run, rsim_copy_states, rsim_cost_features / predict / order, rsim_gather_states, rsim_permute_words)
 * are recorded instead of run; a synchronising call fails inside a capture and ends it. rsim_graph_launch replays the
 * recorded work on the stream (asynchronously) with the buffers it was recorded with. No counterpart in the reference. */
typedef struct rsim_graph rsim_graph;
int rsim_capture_begin(rsim_device* dev);
int rsim_capture_end(rsim_device* dev, rsim_graph** out);
int rsim_graph_launch(rsim_device* dev, const rsim_graph* graph);
int rsim_graph_free(rsim_graph* graph);

/* Timing on the device's stream: HIP events recorded before and after the work queued between begin and end. */
int rsim_timer_begin(rsim_device* dev);
int rsim_timer_end(rsim_device* dev, float* ms); /* synchronises; ms = elapsed device time */
int rsim_sync(rsim_device* dev);

/* The message of the calling thread's last failed call ("" if none). */
const char* rsim_last_error(void);

/* ---- synthetic workloads (host only; the stand-in for RocketSim V3's exports, crates/rocketsim-cuda/src/dump.rs) ---- */
#define RSIM_NUM_PRESETS 7 /* Octane, Dominus, Plank, Breakout, Hybrid, Merc, Psyclops (v3dump/src/scenario.rs:21-29) */
#define RSIM_SCEN_KICKOFF 0   /* one of the 5 kickoff spots, both cars mirrored */
#define RSIM_SCEN_OPEN 1      /* open play: cars and ball anywhere, random pads */
#define RSIM_SCEN_PINCH 2     /* the ball within 200 uu of a wall or fillet, a car driving into it */
#define RSIM_SCEN_HEADON 3    /* two cars closing at 1500-2300 uu/s from 300-1200 uu */
#define RSIM_SCEN_GOALFRAME 4 /* the ball within 150 uu of a post or the crossbar, moving at it */
#define RSIM_SCEN_CONTACT 5   /* 40% pinch, 30% head-on, 30% goal frame */
#define RSIM_SCEN_MIXED 6     /* 1:1:1 kickoff, open play, contact */
#define RSIM_SCEN_OPEN_CONTACT 7 /* 1:1 open play, contact */

/* The StaticData of one world: the procedural Soccar arena, car 0 with preset0 and car 1 with preset1, and the mutators
 * (Soccar defaults: gravity_z -650, boost_accel_scale 1, ball_radius 91.25). */
int rsim_build_world(int preset0, int preset1, float gravity_z, float boost_accel_scale, float ball_radius, void* world_out);
/* Arenas first..first+n-1 of a seeded workload: arena i depends only on (seed, i), so any subrange is reproducible. */
int rsim_sample_arenas(const void* world, int scenario, uint64_t seed, uint32_t first, uint32_t n, void* states_out);
/* As rsim_sample_arenas, with arena first + i drawn in worlds[variant[i]] (a host table of `count` StaticData, e.g. the
 * one passed to rsim_upload_worlds); arena i is the same draw rsim_sample_arenas makes in that world. */
int rsim_sample_arenas_in(const void* worlds, uint32_t count, const uint32_t* variant, int scenario, uint64_t seed,
                          uint32_t first, uint32_t n, void* states_out);
/* The controls of one decision for arenas first..first+n-1 (Controls[n][2]): axes uniform in {-1, 0, 1}, jump / boost /
 * handbrake Bernoulli(0.3). */
int rsim_sample_controls(uint64_t seed, uint32_t decision, uint32_t first, uint32_t n, void* controls_out);

#ifdef __cplusplus
}
#endif
#endif /* RSIM_H */
