/*
 * adfwi_amd.h -- C ABI of libadfwi_amd.so, the MI355X (gfx950) propagator
 * library for ADFWI's acoustic and elastic staggered-grid time stepping.
 *
 * Conventions
 *   - Every buffer is caller-owned device memory; the library never allocates
 *     or frees device memory and keeps no mutable global state, so calls on
 *     distinct streams may run concurrently.  The one exception is
 *     adfwi_acoustic_adjoint_noalloc, which uses a per-GPU scratch array of the
 *     library (see its declaration).
 *   - Functions return 0 on success or a negative ADFWI_E* code;
 *     adfwi_last_error() returns a thread-local message for the last failure
 *     on the calling thread.  Arguments are checked before any launch.
 *   - Grids are the padded grids of the reference (absorbing layers and, for
 *     the elastic case, the two free-surface rows included).  Every 2-D field
 *     is stored row-major with a row pitch of `pitch` floats (pitch >= nx);
 *     a state set is [nfields][nshots][nz][pitch].
 *   - Streams are passed as opaque pointers (a hipStream_t; NULL = default).
 *
 * The entry points replace the reference's per-step PyTorch op sequences:
 *   acoustic : ADFWI/propagator/acoustic_kernels.py:40-176 (step_forward) and
 *              its autograd replay under torch.utils.checkpoint (:267-278);
 *   elastic  : ADFWI/propagator/elastic_kernels.py:222-423
 *              (step_forward_PML_4order) and its autograd replay (:975-992).
 */
#ifndef ADFWI_AMD_H
#define ADFWI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ADFWI_ABI_VERSION 3

#define ADFWI_OK 0
#define ADFWI_E_VERSION -1   /* abi_version / struct_size mismatch */
#define ADFWI_E_ARG -2       /* invalid argument (shape, index, null pointer) */
#define ADFWI_E_LAUNCH -3    /* HIP launch or runtime failure */
#define ADFWI_E_FAULT_INJECTED -4 /* ADFWI_FAULT_INJECT was set (testing) */

typedef void *adfwi_stream_t;

/* Version of this ABI (ADFWI_ABI_VERSION). */
int adfwi_abi_version(void);
/* Static text for an error code. */
const char *adfwi_strerror(int code);
/* Thread-local description of the last failure on this thread. */
const char *adfwi_last_error(void);

/* ------------------------------------------------------------------------ */
/* Acoustic p-u-w propagator (acoustic_kernels.py:113-174)                   */
/* ------------------------------------------------------------------------ */

typedef struct adfwi_acoustic_desc {
  int32_t abi_version;     /* ADFWI_ABI_VERSION */
  int32_t struct_size;     /* sizeof(adfwi_acoustic_desc) */
  int32_t nshots;          /* S: shots advanced together */
  int32_t nz, nx;          /* padded grid Z = nz+2*nabc, X = nx+2*nabc */
  int32_t pitch;           /* row pitch of every field, in floats (>= nx) */
  int32_t fs;              /* first free-surface row: nabc with free surface, else 1 */
  int32_t free_surface;    /* 0 / 1 */
  int32_t nt;              /* time steps of the whole record */
  int32_t nrec;            /* R: receivers shared by all shots */
  int32_t shots_per_block; /* shots advanced by one wavefront: 1, 2, 4, 8; 0 = library choice */
  int32_t row_unroll;      /* row-loop unroll of the step kernels: 1, 2, 4; 0 = library choice */
  int32_t tile_z;          /* rows of one strip: 4..64; 0 = library choice */
  int32_t adj_shots_per_block; /* shots per wavefront of the two-step adjoint kernel: 1, 2, 4;
                                  0 = library choice */
  const float *coef;       /* [5][nz][pitch]: 1-K1, A1, 1-K2, A2, 1-K3 */
  const float *src;        /* [nshots][nt]: dt * wavelet, injected into p */
  const int32_t *src_z;    /* [nshots] padded source rows */
  const int32_t *src_x;    /* [nshots] padded source columns */
  const int32_t *rcv_z;    /* [nrec] padded receiver rows (< nz-1) */
  const int32_t *rcv_x;    /* [nrec] padded receiver columns (< nx-1) */
  const int32_t *coef_set; /* [nshots] or NULL: index of each shot's coefficient set, the sets
                              stored one after another in `coef` ([nsets][5][nz][pitch]), so
                              that shots of requests against different models share launches.
                              Every aligned group of max(layout.shots_per_block, layout.reserved)
                              shots must use one set.  No counterpart in the reference (one model
                              per propagator). */
} adfwi_acoustic_desc;

typedef struct adfwi_layout {
  int32_t tile_z, tile_x;  /* output tile (strip) of one wavefront */
  int32_t zbase;           /* first grid row covered by tiles (rows above never change) */
  int32_t ntiles_z, ntiles_x, ntiles;
  int32_t shots_per_block; /* effective value */
  int32_t ngroups;         /* shot groups = ceil(nshots / shots_per_block) */
  int32_t nfields;         /* fields in one state set */
  int32_t reserved;        /* acoustic: shots per wavefront of the two-step adjoint kernel */
  int64_t field_floats;    /* nshots * nz * pitch */
  int64_t state_floats;    /* nfields * field_floats */
  int64_t accum_floats;    /* acoustic: one accumulator plane (nz * pitch) per shot group of the
                              finer of the two groupings (shots_per_block, reserved) */
} adfwi_layout;

/* Tiling and buffer sizes for a descriptor (host only, no device access). */
int adfwi_acoustic_get_layout(const adfwi_acoustic_desc *d, adfwi_layout *out);

typedef struct adfwi_acoustic_fwd_args {
  const float *state_in;   /* [3][S][Z][pitch] p,u,w before step t0 */
  float *work[2];          /* two scratch state sets */
  float *ckpt;             /* [n][3][S][Z][pitch] or NULL: slot j <- state after step (j+1)*ckpt_every-1 */
  int32_t ckpt_every;
  int32_t t0, nsteps;      /* advance steps t0 .. t0+nsteps-1 */
  int32_t energy_u_from;   /* first step whose u^2 enters the u energy map */
  float *rec_p, *rec_u, *rec_w; /* [S][nt][R] or NULL */
  float *energy;           /* [ngroups][2][Z][pitch] (+=) or NULL: sum over shots of p^2 every
                              step and of u^2 from step energy_u_from on */
  float *image;            /* [nsteps][S][Z][pitch] or NULL: p-update stencil of each step */
  float *image_p;          /* [nsteps][S][Z][pitch] or NULL: p after each step */
  const float *state_out;  /* OUT: buffer holding the state after the last step */
  const int32_t *rcv_map;  /* [nz + nz*pitch] or NULL: receivers per row, then the receiver index
                              at each point (-1 elsewhere); lets two steps share one launch while
                              records are taken (receivers distinct, none on row fs-1 with a free
                              surface) */
  int32_t fuse_steps;      /* 1: advance two steps per launch where possible; 0: one */
  int32_t reserved2;
  const int32_t *t_base;   /* device int or NULL: added at run time to every step index of the
                              call (t0 .. t0+nsteps-1 are then relative), so that one captured
                              launch sequence (a HIP graph) serves every segment of a run; the
                              caller keeps *t_base + t0 + nsteps <= nt (steps outside [0, nt)
                              are skipped).  Requires ckpt == NULL. */
} adfwi_acoustic_fwd_args;

/* Forward time stepping (replaces step_forward, acoustic_kernels.py:40-176). */
int adfwi_acoustic_forward(const adfwi_acoustic_desc *d, adfwi_acoustic_fwd_args *a,
                           adfwi_stream_t stream);

typedef struct adfwi_acoustic_adj_args {
  const float *lam_in;     /* adjoint state after step t_hi, receiver sources of t_hi not yet
                              added; must be work[0] or work[1] (sources are added in place) */
  float *work[2];
  int32_t t_hi, nsteps;    /* adjoints of steps t_hi, t_hi-1, ..., t_hi-nsteps+1 */
  int32_t image_t0;        /* image slot k holds step image_t0 + k */
  int32_t reserved;
  const float *image;      /* stencil images written by adfwi_acoustic_forward */
  const float *image_p;    /* pressure images, or NULL when grad_a2 is NULL */
  const float *grad_rec_p, *grad_rec_u, *grad_rec_w; /* [S][nt][R] or NULL */
  float *grad_a1, *grad_a2; /* [ngroups][Z][pitch] (+=) or NULL */
  float *grad_src;          /* [S][nt] or NULL: d loss / d src */
  const float *lam_out;     /* OUT: adjoint state after step t_hi-nsteps (its sources not added) */
  const int32_t *rcv_map;   /* as in adfwi_acoustic_fwd_args: lets two steps share one launch
                               while record adjoints are injected */
  int32_t fuse_steps;       /* 1: reverse two steps per launch where possible; 0: one */
  int32_t reserved2;
} adfwi_acoustic_adj_args;

/* Discrete adjoint of adfwi_acoustic_forward (the transpose of every stage of
 * acoustic_kernels.py:115-169), with zero-lag imaging of dL/dA1 and dL/dA2. */
int adfwi_acoustic_adjoint(const adfwi_acoustic_desc *d, adfwi_acoustic_adj_args *a,
                           adfwi_stream_t stream);

/* Size of the library's per-GPU scratch array used by adfwi_acoustic_adjoint_noalloc, in
 * floats (128 MiB). */
#define ADFWI_NOALLOC_SCRATCH_FLOATS (32ll << 20)

/* Allocation-free variant of adfwi_acoustic_adjoint.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same kernels run; the per-shot-group gradient
 * planes live in a file-scope device array of the library (ADFWI_NOALLOC_SCRATCH_FLOATS floats
 * per GPU, held until the process ends), are cleared by a fill kernel at the start of every
 * call and summed in plane order into the caller's one-plane outputs at its end:
 *   grad_a1_sum, grad_a2_sum: [nz][pitch] (+=) or NULL (then the gradient is not formed);
 *   a->grad_a1 and a->grad_a2 must be NULL; every other field of `a` as for
 *   adfwi_acoustic_adjoint (the caller still passes inputs, outputs and work state sets).
 * Fails with ADFWI_E_ARG when the planes (layout.accum_floats per requested gradient) exceed
 * the scratch array.  Calls no hipMalloc/hipFree.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race. */
int adfwi_acoustic_adjoint_noalloc(const adfwi_acoustic_desc *d, adfwi_acoustic_adj_args *a,
                                   float *grad_a1_sum, float *grad_a2_sum, adfwi_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Elastic velocity-stress propagator, 4th order, split-field PML             */
/* (elastic_kernels.py:222-423 step, :917-1037 driver)                        */
/* ------------------------------------------------------------------------ */

typedef struct adfwi_elastic_desc {
  int32_t abi_version;     /* ADFWI_ABI_VERSION */
  int32_t struct_size;     /* sizeof(adfwi_elastic_desc) */
  int32_t nshots;          /* S */
  int32_t nz, nx;          /* padded grid (the nn free-surface padding rows on top included) */
  int32_t pitch;           /* row pitch in floats (>= nx) */
  int32_t free_surface;    /* 0 / 1: free-surface rows nn-2 .. nn+1 as in the reference */
  int32_t nt, nrec;
  int32_t aniso;           /* 1 when C15 or C35 is non-zero (general 2-D anisotropy) */
  int32_t shots_per_wave;  /* 0 = library choice */
  int32_t tile_z;          /* strip rows, 0 = library choice */
  float dt, dt_dx, dt_dz, dx, dz;  /* fp32 scalars as the reference rounds them */
  float c0, c1;            /* staggered FD coefficients DiffCoef(2, 's') */
  int32_t damping;         /* 0: split-field PML (step_forward_PML_4order, ten split fields);
                              1: multiplicative damping (step_forward_ABL_4order, elastic_kernels.py
                              :585-777: five fields txx, tzz, txz, vx, vz, -MT/3 source) */
  const float *coef;       /* [8][nz][pitch]: C11, C13, C33, C55, bx, bz, C15, C35 (padded as in
                              the reference, each from its own shape; row/col 0 based) */
  const float *pml_x;      /* PML: [3][pitch]: 1 - dt/2 BCx, 1 + dt/2 BCx, 1/(1 + dt/2 BCx) per column */
  const float *pml_z;      /* PML: [3][nz]: the same for BCz per row */
  const float *src;        /* [nshots][nt] wavelets */
  const float *src_m;      /* [nshots][3]: PML -MT_xx/2, -MT_zz/2, -MT_xz/2; damping -MT/3 likewise */
  const int32_t *src_z, *src_x;   /* [nshots] padded source indices (inside the window) */
  const int32_t *rcv_z, *rcv_x;   /* [nrec] padded receiver indices (inside the window) */
  const float *damp;       /* damping: [nz][pitch] multiplicative taper of vx, vz; PML: unused */
  const float *src_fz;     /* [nshots] or NULL: vertical-force weight f of each shot's wavelet w,
                              added to vz at the source point right after the velocity update
                              (PML: f w / 2 into each vz split; damping: f w, before the taper).
                              No counterpart in the reference, whose sources are moment tensors. */
  int32_t nn;              /* stencil half-order NN = fd_order / 2: 0 or 2 for the 4th-order schemes;
                              3, 4, 5 for the three-term PML scheme of fd_order 6, 8, 10
                              (step_forward_PML_6order, elastic_kernels.py:426-584: window rows
                              [NN, nz-NN-1], free-surface rows NN-2 .. NN+1, padding NN rows on top) */
  float c2;                /* third coefficient of DiffCoef(NN, 's') (NN >= 3) */
  const int32_t *rcv_map;  /* [nz + nz*pitch] or NULL: receivers per row, then the receiver index at
                              each point (-1 elsewhere; every receiver at its own point).  With it the
                              4th-order PML step kernels take the records and add the record
                              adjoints themselves (elastic_kernels.py:407-411 samples inside the
                              step too); without it, or with another scheme, separate launches do. */
} adfwi_elastic_desc;

/* Tiling and buffer sizes; layout.nfields = 10 (PML split fields) or 5 (damping), layout.reserved = shots per
 * wavefront, layout.shots_per_block = shots per workgroup, layout.ngroups = workgroups along
 * the shots, layout.accum_floats = size of the gradient accumulator of the adjoint,
 * [ngroups][8][nz][pitch] (one set of planes per workgroup's shots). */
int adfwi_elastic_get_layout(const adfwi_elastic_desc *d, adfwi_layout *out);

typedef struct adfwi_elastic_fwd_args {
  const float *state_in;   /* [nfields][S][Z][pitch] state before step t0 */
  float *work[2];
  float *ckpt;             /* [n][10][S][Z][pitch] or NULL */
  int32_t ckpt_every;
  int32_t t0, nsteps;
  int32_t reserved;
  float *rec[5];           /* txx, tzz, txz, vx, vz records [S][nt][R], or all NULL */
  float *trace;            /* [nsteps][10][S][Z][pitch] or NULL: the state after every step of the
                              call is written to its slot (work[] unused; ckpt must be NULL) */
  const float *state_out;  /* OUT */
} adfwi_elastic_fwd_args;

/* Forward time stepping (replaces step_forward_PML_4order, elastic_kernels.py:222-423,
 * step_forward_PML_6order :426-584 and step_forward_ABL_4order :585-777). */
int adfwi_elastic_forward(const adfwi_elastic_desc *d, adfwi_elastic_fwd_args *a, adfwi_stream_t stream);

typedef struct adfwi_elastic_adj_args {
  float *lam_in;           /* adjoint of the state after step t_hi-1: work[0] or work[1]; without
                              rcv_map the record adjoints of each reversed step are added to it in
                              place (with it, inside the step kernels: lam_in is then only read) */
  float *work[2];          /* [10][S][Z][pitch] each, distinct */
  int32_t t_hi;            /* steps t_hi-1 down to t_hi-nsteps are reversed */
  int32_t nsteps;
  const float *fwd0;       /* forward state before step t_hi-nsteps [10][S][Z][pitch] */
  const float *trace;      /* forward states after steps t_hi-nsteps .. t_hi-1 [nsteps][10][S][Z][pitch] */
  const float *grad_rec[5];/* dL/d(record) [S][nt][R] for txx, tzz, txz, vx, vz; NULL = zero */
  float *grad_coef;        /* accumulated: layout.accum_floats, planes C11, C13, C33, C55, bx, bz,
                              C15, C35 of every workgroup's shots (sum over the groups) */
  float *grad_src;         /* [S][nt] or NULL: dL/d(wavelet) written for the reversed steps */
  const float *lam_out;    /* OUT: adjoint of the state before step t_hi-nsteps */
} adfwi_elastic_adj_args;

/* Reverse time stepping: the exact transpose of adfwi_elastic_forward's steps, with the
 * zero-lag cross-correlation gradients of the padded coefficient fields.  Replaces the
 * autograd replay of step_forward_PML_4order (elastic_kernels.py:975-1008 checkpoints). */
int adfwi_elastic_adjoint(const adfwi_elastic_desc *d, adfwi_elastic_adj_args *a, adfwi_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Record post-processing                                                     */
/* ------------------------------------------------------------------------ */

#define ADFWI_L2_NORMALIZE 1  /* n = a / max_t |a| per trace (fwi/acoustic_fwi.py:149-150) */
#define ADFWI_L2_SQUARED 2    /* least-squares trace loss (no counterpart in the reference) */

/* Waveform L2 misfit per trace of records a against b ([nshots][nt][nrec], device):
 * trace_loss[s][r] = sqrt(sum_t (n - b)^2 dt) (Misfit_waveform_L2, fwi/misfit/L2.py:26-29), or
 * 1/2 sum_t (n - b)^2 dt with ADFWI_L2_SQUARED (the least-squares objective, whose gradient is
 * defined on traces with zero residual, where the square root's is NaN as in the reference);
 * n = a / max_t |a| per trace with ADFWI_L2_NORMALIZE (the FWI drivers' trace normalisation),
 * else n = a.  The misfit is the sum of trace_loss.  grad_a / grad_b (or NULL): d misfit / d a
 * and d misfit / d b, through the normalisation (whose max routes its gradient to the first
 * sample reaching it).  flags = 0 or 1 keeps the meaning of the earlier normalize_a argument. */
int adfwi_trace_l2(const float *a, const float *b, int32_t nshots, int32_t nt, int32_t nrec, float dt,
                   int32_t flags, float *trace_loss, float *grad_a, float *grad_b, adfwi_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ADFWI_AMD_H */
