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
} jp_bench_result;

/* Creates a uniform context and allocates its two population buffers (src/tgv3d_uniform.cu:157-159).
 * Returns JP_ERR_ARG when n % 8 != 0 or n < 8 (src/tgv3d_uniform.cu:148). */
int jp_uniform_create(const jp_uniform_params* p, jp_ctx** out);

/* TGV initial state, populations at equilibrium (src/tgv3d_uniform.cu:161 k_init_tgv3d; for AMR also
 * the initial build of src/tgv3d_amr.cu:776-815). */
int jp_init(jp_ctx* c);

/* n time steps: uniform k_step + swap (src/tgv3d_uniform.cu:194); AMR: n coarse steps of the nested
 * 1:2:4 subcycle base_step (src/tgv3d_amr.cu:821-837). */
int jp_step(jp_ctx* c, long n);

/* Energy <|u|^2/2>, enstrophy <|omega|^2/2>, mass <rho>, reduced on the device in FP64 with a fixed
 * order; 24 bytes cross to the host (src/tgv3d_uniform.cu:170-183; src/tgv3d_amr.cu:882-885). */
int jp_diag(jp_ctx* c, double out[3]);

/* Pure step loop timed with HIP events on the context's stream (src/tgv3d_amr.cu:839-864; the uniform
 * reference has no bench path). use_graph replays the steps from a captured hipGraph. */
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

/* Library version string. */
const char* jp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* JP_H */
