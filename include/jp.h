/* jp.h - C ABI of the MI355X (gfx950) D3Q19 Taylor-Green time step.
 *
 * The reference (josepedro/jp-solver) has no library interface: its hot path is a set of static
 * __global__ kernels launched from main() of src/tgv3d_uniform.cu and src/tgv3d_amr.cu. Each entry
 * point below replaces one of those in-process seams; the reference lines it stands for are cited.
 *
 * Conventions
 *  - Every call returns JP_OK (0) or a JP_ERR_* code; jp_last_error() gives the message. No call exits
 *    the process (the reference's CK() exits with status 2, src/tgv3d_uniform.cu:25-26).
 *  - A context owns all device memory, its HIP stream and its graphs. Host buffers passed in are owned
 *    by the caller and never retained.
 *  - A context is used from one host thread at a time; independent contexts may live on different
 *    threads.
 *  - Inside jp_step / jp_regrid nothing crosses device-to-host except scalar counts.
 *  - Population buffers crossing this boundary use the reference's tile-major layout
 *    f[(b*19 + q)*512 + cell], cell = (k*8 + j)*8 + i, b = (bz*gx + by)*gx + bx
 *    (src/tgv3d_uniform.cu:42-53), whatever the layout on the device.
 */
#ifndef JP_H
#define JP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JP_OK 0
#define JP_ERR_HIP 1      /* HIP runtime error (reference: CK() -> exit 2) */
#define JP_ERR_ARG 2      /* bad argument, e.g. N % 8 != 0 (reference: src/tgv3d_uniform.cu:148) */
#define JP_ERR_CAPACITY 3 /* a refinement count exceeded its pool (the reference writes out of bounds) */
#define JP_ERR_STATE 4    /* call not valid for this context or state */
#define JP_ERR_NOMEM 5    /* device allocation failed */

/* Population storage forms. FP16/BF16 store (f - w_q) * scale with a power-of-two scale per level and
 * collide in FP32; FP64 stores and collides in double. FP32 is the reference's form. */
#define JP_PREC_FP32 0
#define JP_PREC_FP16 1
#define JP_PREC_BF16 2
#define JP_PREC_FP64 3

typedef struct jp_ctx jp_ctx;

/* Uniform lattice parameters: the ./uniform flags of src/tgv3d_uniform.cu:143-147. */
typedef struct {
  int n;           /* --n: lattice edge, a multiple of 8 */
  double re;       /* --re */
  double u0;       /* --u0: lattice velocity amplitude */
  int prec;        /* JP_PREC_* */
  double phase[3]; /* TGV phase shift per axis in radians (0 = the reference's initial state) */
  double scale;    /* FP16/BF16 store scale; 0 = default 2^floor(log2(2048/u0)) */
  int device;      /* HIP device ordinal */
  long plane_pad;  /* device population layout: 0 = the library's default (z-layers, the 19 directions
                      of one z-layer adjacent; direction planes for FP64); -1 = 19 direction planes; a
                      positive multiple of 64 = direction planes with that many elements appended to
                      each (HBM channel spread). Every layout computes the same result. */
} jp_uniform_params;

/* Derived run constants: src/tgv3d_uniform.cu:151-153 (uniform) and src/tgv3d_amr.cu:659-664 (AMR). */
typedef struct {
  double nu, tau0, tau1, tau2; /* viscosity, relaxation time per level (tau1, tau2 = 0 for uniform) */
  double omega0, omega1, omega2;
  double kf;                   /* 2 pi / N */
  double nqs01, nqi10, nqs12, nqi21;
  double scale[3];             /* store scale per level */
  int prec;
  int n;
  long cells;                  /* cells of level 0 (N^3) */
} jp_derived;

typedef struct {
  double ms_total;      /* event time of the timed steps */
  double ms_per_step;
  double work;          /* cell updates in the timed region, fine levels weighted 2 and 4 */
  double glups;         /* work / time, 1e9 updates per second */
  double bytes_per_update;
  double gbps;          /* glups * bytes_per_update */
  double pool_gb;       /* device bytes held by the context */
  long steps;
  double device_used_gb; /* device memory in use after the warm-up, total - free (hipMemGetInfo): the
                            reference's "ctx+pools" figure (src/tgv3d_amr.cu:841, 861-862) */
} jp_bench_result;

/* Refinement modes of --l1 / --l2 (src/tgv3d_amr.cu:651, 777-797). */
#define JP_MODE_ALL 0
#define JP_MODE_SLAB 1
#define JP_MODE_SENSOR 2

/* 3-level block AMR parameters: the ./amr flags of src/tgv3d_amr.cu:644-656, 674, 685. */
typedef struct {
  int n;             /* --n: coarse edge N0 (multiple of 8); the finest level is 4 N0 */
  double re, u0;     /* --re, --u0 */
  int prec;          /* JP_PREC_* */
  double phase[3];   /* TGV phase shift per axis (0 = reference) */
  int l1_mode;       /* --l1: JP_MODE_* (used at the initial build only, as in the reference) */
  int l2_mode;       /* --l2 */
  double frac1;      /* --frac1: target refined fraction of level 0 in (0, 1] */
  double frac2;      /* --frac2: target refined fraction of the deep level-1 tiles in (0, 1] */
  int adapt;         /* --adapt: regrid period in coarse steps (used by jp_bench and the driver) */
  double maxgb;      /* --maxgb: fine-pool budget in GiB, 0 = 75 % of free device memory */
  int skipprep;      /* --skipprep: profiling only, skips prep + ghost fill (breaks the physics) */
  int skiprestr;     /* --skiprestr: profiling only, skips restriction (breaks the physics) */
  double scale[3];   /* FP16/BF16 store scale per level; 0 = default */
  int device;
} jp_amr_params;

/* Refinement counts and pool capacities (src/tgv3d_amr.cu:681-683, 708, 735, 752, 759-761). */
typedef struct {
  long cnb0, nref0, nf1, nref1, nf2, n_skin0, n_skin1;
  long cap_nref0, cap1, cap_nref1, cap2, hw_nf1, hw_nf2, ndeep;
} jp_counts;

/* Table ids for jp_get_table (int32 tables, except the two float sensors). */
#define JP_TABLE_REF0 0
#define JP_TABLE_FOF0 1
#define JP_TABLE_RL0 2
#define JP_TABLE_L1PCB 3
#define JP_TABLE_L1OCT 4
#define JP_TABLE_F1POS 5
#define JP_TABLE_L1GRID 6
#define JP_TABLE_REF1 7
#define JP_TABLE_FOF1 8
#define JP_TABLE_RL1 9
#define JP_TABLE_L2PL1 10
#define JP_TABLE_L2OCT 11
#define JP_TABLE_SKIN0 12
#define JP_TABLE_SKIN1 13
#define JP_TABLE_SEN0 14
#define JP_TABLE_SEN1 15

/* Creates a uniform context and allocates its two population buffers (src/tgv3d_uniform.cu:157-159).
 * Returns JP_ERR_ARG when n % 8 != 0 or n < 8 (src/tgv3d_uniform.cu:148). */
int jp_uniform_create(const jp_uniform_params* p, jp_ctx** out);

/* Creates an AMR context: level parameters and pool sizing against the device-memory budget
 * (src/tgv3d_amr.cu:657-709). The level-2 pool is sized by jp_init, after the first measurement. */
int jp_amr_create(const jp_amr_params* p, jp_ctx** out);

/* AMR regrid: sensor, exact threshold, marks, scans, block tables, migration, skins, on the device;
 * only counts come back (src/tgv3d_amr.cu:722-764). JP_ERR_CAPACITY if a count exceeds its pool.
 * With graphs on (jp_use_graph) the regrid of a sensor-driven context runs on device-resident counts:
 * nothing is read back and the host does not wait; a count beyond a pool is reported as JP_ERR_CAPACITY
 * by the next call that reads the counts (jp_get_counts, jp_get_table, jp_diag, the population calls). */
int jp_regrid(jp_ctx* c);

/* Test hooks: the initial build / a regrid with the sensors replaced by the given values, so that two
 * implementations can be fed the same sensor field. sen0 has one value per level-0 tile; sen1_by_pos one
 * value per level-1 grid position ((2N0/8)^3 entries), looked up through each level-1 tile's position.
 * Either may be NULL (computed sensor). */
int jp_init_with_sensor(jp_ctx* c, const float* sen0, const float* sen1_by_pos);
int jp_regrid_with_sensor(jp_ctx* c, const float* sen0, const float* sen1_by_pos);

/* One member of a batched ensemble (not in the reference). */
typedef struct {
  int n;            /* lattice edge (uniform) or coarse edge N0 (AMR); multiple of 8 */
  double re, u0;
  double phase[3];  /* TGV phase shift per axis */
  int amr;          /* 0 uniform lattice, 1 three-level AMR */
  int l1_mode, l2_mode;
  double frac1, frac2;
  int adapt;        /* regrid period of a sensor-driven AMR member in jp_bench */
} jp_member;

/* Batched ensemble (not in the reference; built for many small lattices whose single launches would
 * leave the GPU idle): `count` independent TGV lattices, uniform and AMR mixed, sharing one stream. One
 * jp_step advances every member by one step (uniform) or one coarse step (AMR, 1:2:4 subcycle) with one
 * persistent launch per substep kind over all members' tiles; jp_use_graph captures it in a hipGraph;
 * jp_regrid regrids every sensor-driven AMR member; jp_bench reports the aggregate subcycle-weighted
 * throughput. Members are reachable with jp_ensemble_member for diagnostics and the test hooks. */
int jp_ensemble_create(const jp_member* members, int count, int prec, double maxgb, int device, jp_ctx** out);
/* Tile shape of the batched dense step (not in the reference; a tuning knob that jp_autotune also sets):
 * set = 1 or 8 cells per lane (8 needs uniform members only, each lattice a multiple of 2048 cells), 0 to only
 * query; *current receives the shape in effect. Every shape computes the same result. */
int jp_ensemble_dense_shape(jp_ctx* ens, int set, int* current);
/* Borrowed handle of member `index` (owned by the ensemble, valid until jp_destroy of the ensemble). */
int jp_ensemble_member(jp_ctx* ens, int index, jp_ctx** member);
/* Autotuner (not in the reference; the reference has fixed launch shapes). Candidates are timed on the
 * current state; each timed kernel is re-run on unchanged inputs, so the state is unchanged.
 * - Ensemble: picks the workgroup count of every batched launch, and the dense step's tile shape (1 or 8
 *   cells per lane, jp_ensemble_dense_shape); result in groups[JP_ENS_KERNELS] (dense step, level-1
 *   fill, level-2 fill, level-2 step, level-2 step + restriction, level-1 step, level-1 step +
 *   restriction). The two launches with a restriction write the parent level, so they are not timed:
 *   they always run one workgroup per tile, which overlaps each tile's restriction tail.
 * - Uniform lattice: picks how steps run: single steps in one of 18 launch variants (jp_set_variant: 256,
 *   512 or 1024 lanes per workgroup, 1, 2 or 8 cells per lane, streaming or cached stores), or fused
 *   pairs with one of the column shapes that divide the edge; keeps the fastest, and the fastest single
 *   variant for the odd step after pairs. groups[0..3] receive the chosen column shape (x, y, z extent,
 *   lanes; zeros for single steps), groups[4] the single-step variant, groups[5] the waves per CU the kept
 *   kernel runs at (workgroups a CU holds at once, from its registers, LDS and size, x waves per
 *   workgroup), groups[6] the steps per launch (1 or 2).
 * An AMR context has no candidates (JP_ERR_STATE). */
#define JP_ENS_KERNELS 7
int jp_autotune(jp_ctx* c, int* groups);

int jp_get_counts(jp_ctx* c, jp_counts* out);
/* Copies table `which` (JP_TABLE_*) to dst (capacity cap elements); *n receives its length. */
int jp_get_table(jp_ctx* c, int which, void* dst, long cap, long* n);

/* TGV initial state, populations at equilibrium (src/tgv3d_uniform.cu:161 k_init_tgv3d; for AMR also
 * the initial build of src/tgv3d_amr.cu:776-815). */
int jp_init(jp_ctx* c);

/* n time steps: uniform k_step + swap (src/tgv3d_uniform.cu:194); AMR: n coarse steps of the nested
 * 1:2:4 subcycle base_step (src/tgv3d_amr.cu:821-837). */
int jp_step(jp_ctx* c, long n);

/* When on, jp_step replays pairs of steps from captured hipGraphs, as the reference's --graph
 * production loop does (src/tgv3d_amr.cu:872-881, 896). Pairs are captured once per role of the
 * ping-pong buffers and kept. For a sensor-driven AMR context, turning graphs on also sizes the fine
 * pools for the largest counts a regrid can produce (within the budget) and switches jp_regrid to
 * device-resident counts: every level-1 and level-2 launch reads its tile count on the device, so the
 * captured pairs stay valid across regrids and are not recaptured (the reference recaptures after each
 * regrid). */
int jp_use_graph(jp_ctx* c, int on);

/* Launch variant of the dense-lattice step (not in the reference; a tuning knob): bits 0-1 select the
 * workgroup size (256, 128, 512, 1024 lanes), bit 2 cached instead of streaming population loads, bit 3
 * cached instead of streaming stores, bits 4-5 consecutive x-cells per lane (1, 2, 4, 8) read and written
 * as one vector per direction. Every variant computes the same result. */
int jp_set_variant(jp_ctx* c, int variant);

/* Two fused time steps per launch for the uniform lattice (not in the reference, which launches one
 * k_step per step, src/tgv3d_uniform.cu:166-167; a tuning knob): a workgroup streams a column of the
 * lattice and keeps the intermediate step in LDS, so HBM sees one read and one write per two steps plus
 * the column halos. On by default for FP32, FP16 and BF16 storage where a column shape divides N (a
 * multiple of 32); FP64 and other edges run single steps. Results are bit-identical either way; an odd step count ends
 * with a single step. Resets the column shape to the first in preference order that divides N. */
int jp_set_step_pairs(jp_ctx* c, int on);
/* Selects the fused pair's column shape by its index in the library's preference order (0: 128 x 4,
 * 1: 64 x 8, 2: 32 x 16 cells, each with the longest z-segment of 256, 128, 64 or 32 planes that divides
 * N; 3-5: the same columns with half that segment, at least 32 planes; -1: the first that applies) and
 * turns the pairs on (not in the
 * reference; a tuning knob, as jp_set_variant). JP_ERR_ARG when the shape does not divide N or the
 * storage form does not fuse (FP64); JP_ERR_STATE on an AMR context. Every shape computes the same
 * result, in every storage form. */
int jp_set_step_pair_shape(jp_ctx* c, int pick);
/* The fused pair's column shape {x, y, z extent, lanes per workgroup}; all zero when steps run singly. */
int jp_step_pair_shape(const jp_ctx* c, int shape[4]);
/* Steps one launch of the uniform lattice advances: 1 (single steps) or 2 (fused pairs); 1 on AMR and
 * ensemble contexts (not in the reference; reports the tuning knob above). */
int jp_steps_per_launch(const jp_ctx* c, int* steps);

/* Energy <|u|^2/2>, enstrophy <|omega|^2/2>, mass <rho>, reduced on the device in FP64 with a fixed
 * order; 24 bytes cross to the host (src/tgv3d_uniform.cu:170-183; src/tgv3d_amr.cu:882-885). */
int jp_diag(jp_ctx* c, double out[3]);

/* jp_diag for a uniform context without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The partial sums go to a fixed array of the library (768 KiB on
 * every GPU it runs on, held until the process ends), which holds the partials of any lattice
 * jp_uniform_create accepts; a lattice needing more returns JP_ERR_ARG. The result equals jp_diag's bit
 * for bit (same kernels, same summation order). Every call on one GPU shares the array, so calls that
 * overlap on one GPU (contexts on different host threads or streams) race. JP_ERR_STATE for AMR and
 * ensemble contexts. */
int jp_diag_noalloc(jp_ctx* c, double out[3]);

/* Pure step loop timed with HIP events on the context's stream (src/tgv3d_amr.cu:839-864; the uniform
 * reference has no bench path). A sensor-driven AMR context regrids every `adapt` steps in warm-up and
 * timed loop. use_graph replays pairs of steps from captured hipGraphs; the AMR regrids then run on
 * device-resident counts between replays (an ensemble replays without regrids, as the reference's graph
 * bench does). */
int jp_bench(jp_ctx* c, long steps, long warmup, int use_graph, jp_bench_result* r);

int jp_get_derived(jp_ctx* c, jp_derived* d);

/* Test hooks: raw stored populations of a level in the tile-major layout (element size 4, 2, 2, 8 for
 * FP32, FP16, BF16, FP64). bytes must equal the level's size. */
int jp_get_populations(jp_ctx* c, int level, void* dst, size_t bytes);
int jp_set_populations(jp_ctx* c, int level, const void* src, size_t bytes);
/* Size in bytes of a level's population buffer in that layout. */
int jp_population_bytes(jp_ctx* c, int level, size_t* bytes);

/* Synchronises the context's stream. */
int jp_sync(jp_ctx* c);
const char* jp_last_error(const jp_ctx* c);
void jp_destroy(jp_ctx* c);

/* Bandwidth probe: copies `bytes` device-to-device with a plain streaming kernel `iters` times and
 * reports GB/s (read + write). Calibrates the roofline and the HBM traffic counters. */
int jp_probe_copy(int device, size_t bytes, int iters, double* gbps, double* ms_per_iter);
/* Attainable HBM bandwidth of plain kernels (one 16 B element per lane, one pass): mode 0 copy (GB/s
 * of read + write), 1 read only, 2 write only; bytes rounded down to a multiple of 4096. */
int jp_probe_bandwidth(int device, int mode, size_t bytes, int iters, double* gbps);

/* Output field of a frame (src/tgv3d_uniform.cu:96-140; src/tgv3d_amr.cu:551-615): Q-criterion (periodic
 * central differences, FP64), speed and velocity of the uniform lattice, or of the AMR composite at 4 N0,
 * computed on the device; only these arrays cross to the host. edge = jp_field_edge(c) (N or 4 N0); Q and
 * speed hold edge^3 floats, vel 3 edge^3 (or NULL); n3 is the capacity of Q in entries. Index (k*edge+j)*edge+i. */
long jp_field_edge(jp_ctx* c);
int jp_field(jp_ctx* c, float* Q, float* speed, float* vel, long n3);
/* One .vti frame in the reference's format (ImageData, appended raw, UInt64 headers): uniform writes Q, speed
 * and velocity (src/tgv3d_uniform.cu:110-140), AMR Q and speed of the composite (src/tgv3d_amr.cu:602-615). */
int jp_write_vti(jp_ctx* c, const char* path);
/* AMR: the level-1 and level-2 tiles as .vtp line boxes in 4 N0 coordinates, point data "level"
 * (src/tgv3d_amr.cu:617-641). */
int jp_write_boxes(jp_ctx* c, const char* path);

/* Library version string. */
const char* jp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* JP_H */
