/*
 * hfm.h -- C ABI of libhfm.so, the gfx950 stencil + multigrid hot path of
 * the AI-HFM-Solver operator surface.
 *
 * Conventions
 *   - Fields are contiguous row-major (D, H, W) arrays: the reference's
 *     (1, D, H, W, 1) tensors without the unit batch/channel axes.  A 2D field
 *     (1, H, W, 1) is passed as D = 1, a 1D field as D = H = 1.
 *   - Every pointer is device memory; every launcher is asynchronous and
 *     stream-ordered on `stream` (a hipStream_t passed as void*; NULL = the
 *     null stream).
 *   - Every entry point returns 0 on success or a negative HFM_E* code and
 *     never throws.  hfm_last_error() returns a thread-local message.
 *   - Nothing here allocates device memory; all buffers are caller-owned.
 *   - Reference-mode grids: 1 <= D <= 65535, D*H*W <= 2^31 - 1 cells and
 *     H*W < 2^30 cells per plane (a cell is addressed as its plane's base
 *     plus a 32-bit byte offset); other extents return HFM_EINVAL.
 *
 * Reference mode (the reference's Keras layers; FP32):
 *   hfm_conv27_f32          <- Conv3D/Conv2D/Conv1D(k=3, s=1, 'SAME') layers:
 *                              CNN3D_A_n (ART/Anisotropic_3d_Multigrid_7point.ipynb:124-178),
 *                              the 27-point layers of UM/AI_HFM_v3.ipynb cells 13-15
 *                              (CNN3D_central_2nd_dif/xadv/yadv/zadv, CNN3D_Su/Sv/Sw,
 *                              CNN3D_pu/pv/pw), and the 2D/1D twins
 *                              (ARTn/Parallel_multigrid_2D.ipynb#c9, _1D#c8); plus
 *                              the fused residual (A x - b) and smoother
 *                              (w - A(w)/d + r/d, …7point.ipynb:424-444) forms.
 *   hfm_var6_f32            <- A_operator_n (ART/BGS_multigrid_v1.ipynb#c40) and its
 *                              Jacobi sweep with divisor A_MAT_TOT (#c42).
 *   hfm_restrict_mean_f32   <- CNN3D_res_n (…7point.ipynb:194-235).
 *   hfm_resid_restrict_f32  <- "r = CNN3D_A_512(values_p) - b; r_256 = CNN3D_res_512(r)"
 *                              fused (UM/AI_HFM_v3.ipynb:2573-2575).
 *   hfm_prolong_nearest_f32 <- CNN3D_prol_n / UpSampling3D (…7point.ipynb:251-284).
 *   hfm_prolong_jacobi_f32  <- "w_f = prol(w_c); w_f = w_f - A(w_f)/d + r_f/d"
 *                              fused (…7point.ipynb:426-444).
 *   hfm_correct_jacobi_f32  <- "values = values - w; values = values - A(values)/d + b/d"
 *                              fused (…7point.ipynb:448-449).
 *   hfm_bc_f32              <- boundary_condition_top (…_BC.ipynb:340-344) and the
 *                              boundary_condition_velocity/pressure/source/indicator
 *                              plane fills (UM/AI_HFM_v3.ipynb cell 19).
 *   hfm_bluff_body_f32      <- bluff_body: f / (1 + dt*sigma) (UM/AI_HFM_v3.ipynb cell 19).
 *
 * Config mode (anisotropic solve div(sigma grad phi) = b; no reference
 * counterpart beyond the operator form of A_operator_n): work-table driven
 * kernels over any number of instances and levels per launch.
 */
#ifndef HFM_H
#define HFM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFM_OK 0
#define HFM_EINVAL (-1)
#define HFM_EDEVICE (-2)
#define HFM_ENOMEM (-3)

/* storage types of config-mode fields */
#define HFM_STORE_F32 0
#define HFM_STORE_BF16 1
#define HFM_STORE_F16 2   /* IEEE half; the solver scales the fine residual (sexp) */

/* reference-mode stencil forms */
#define HFM_FORM_APPLY 0    /* y = A x                         */
#define HFM_FORM_RESID 1    /* y = A x - b                     */
#define HFM_FORM_JACOBI 2   /* y = x - A(x)/d + b/d            */
#define HFM_FORM_SUBTRACT 3 /* y = b - A x                     */
#define HFM_FORM_AXPY 4     /* y = b + d A x  (explicit stages) */
#define HFM_FORM_JACOBI_MUL 5 /* y = (x - d A x) + d b: d the smoother's reciprocal weight */
#define HFM_FORM_MASK 15
#define HFM_FORM_ZERO_OUTLET 16 /* flag: y = 0 on the x = W-1 plane */

/* flags of hfm_correct_jacobi_f32 / hfm_conv27_sum_f32 */
#define HFM_CORRECT_ZERO_OUTLET 1 /* (x - w) = 0 on the x = W-1 plane before the sweep */
#define HFM_CORRECT_RECIPROCAL 2  /* the sweep in the HFM_FORM_JACOBI_MUL form: d is the
                                     reciprocal weight (constant kernel only) */
#define HFM_SUM_SOURCE_BC 1       /* x = 0 plane takes the value of the x = 1 plane   */

/* boundary fills for hfm_bc_f32 */
#define HFM_BC_TOP_COPY 0      /* plane z=0 <- plane z=1                       */
#define HFM_BC_PRESSURE 1      /* x=W-1 -> 0; other faces copy their neighbour */
#define HFM_BC_SOURCE 2        /* x=0 plane <- x=1 plane                        */
#define HFM_BC_INDICATOR 3     /* all six faces -> 0                            */
#define HFM_BC_ZERO_OUTLET 4   /* x=W-1 plane -> 0 (MG residual / correction)   */

const char* hfm_last_error(void);
int hfm_version(void);
/* sizeof(hfm_mg_task) and sizeof(hfm_mg_scalars) as compiled, for host-side
 * table builders to check their layout against. */
int hfm_struct_sizes(int64_t* task_size, int64_t* scalars_size);

/* ---------------- reference mode ---------------- */

/* y = form(K (*) x): K is the 27-entry row-major 3x3x3 kernel w[0,:,:,:,0];
 * zero ghosts.  b: rhs for RESID/JACOBI/SUBTRACT (may be NULL for APPLY); d:
 * the scalar divisor for JACOBI; form may carry HFM_FORM_ZERO_OUTLET
 * (the MG residual's r[..., nx-1] = 0, UM/AI_HFM_v3.ipynb:2574).  y must not
 * alias x. */
int hfm_conv27_f32(const float* x, float* y, const float* b, const float K[27], float d,
                   int64_t D, int64_t H, int64_t W, int form, void* stream);

/* y = form(A6 x) with A6 x = sum_dir m_dir (x - x_nb) over the six directions
 * (x-, x+, y-, y+, z-, z+ order of `m`), zero ghosts; JACOBI divides by `tot`. */
int hfm_var6_f32(const float* x, float* y, const float* b, const float* const m[6],
                 const float* tot, int64_t D, int64_t H, int64_t W, int form, void* stream);

/* y (D/fz, H/fy, W/fx) = mean of the fz*fy*fx children (factors 1 or 2). */
int hfm_restrict_mean_f32(const float* x, float* y, int64_t D, int64_t H, int64_t W,
                          int fz, int fy, int fx, void* stream);

/* rc (D/2, H/2, W/2) = restrict_mean(K (*) x - b) with factors (2, 2, 2), in
 * one pass and without writing the fine residual: the sawtooth cycle's
 * "r = CNN3D_A_512(values_p) - b; r[..., nx-1] = 0; r_256 = CNN3D_res_512(r)"
 * (UM/AI_HFM_v3.ipynb:2573-2575), for the cycles whose finest level takes no
 * sweep of its own (the fine residual has no other reader).  Bitwise the
 * hfm_conv27_f32 (HFM_FORM_RESID) + hfm_restrict_mean_f32 pair; flags may be
 * HFM_FORM_ZERO_OUTLET.  D, H and W even; rc must not alias x or b. */
int hfm_resid_restrict_f32(const float* x, const float* b, float* rc, const float K[27],
                           int64_t D, int64_t H, int64_t W, int flags, void* stream);

/* y (Dc*fz, Hc*fy, Wc*fx) = nearest-neighbour repeat of x (Dc, Hc, Wc). */
int hfm_prolong_nearest_f32(const float* xc, float* y, int64_t Dc, int64_t Hc, int64_t Wc,
                            int fz, int fy, int fx, void* stream);

/* y (Dc*fz, Hc*fy, Wc*fx) = UpSampling by (fz, fy, fx) followed by the
 * all-ones window-2 convolution with 'SAME' padding on each upsampled axis
 * (zero after the end): the prolongation of ARTn/Multigrid_cnn.ipynb:342-398
 * (UpSampling2D then Conv2D(kernel 2x2 of ones, padding='SAME')). */
int hfm_prolong_upsum_f32(const float* xc, float* y, int64_t Dc, int64_t Hc, int64_t Wc, int fz,
                          int fy, int fx, void* stream);

/* out[0] = sum of the n floats of x in FP64, fixed order (one workgroup): the
 * residual monitor sum(r) / (nx ny) of ARTn/Multigrid_cnn.ipynb:605. */
int hfm_sum_f64(const float* x, int64_t n, double* out, void* stream);

/* y = J(P wc) against r: P = nearest repeat by (fz,fy,fx) onto (D,H,W).
 * Operator: constant kernel K with divisor d (m == NULL) or the var6 fields
 * m/tot (K ignored). */
int hfm_prolong_jacobi_f32(const float* wc, const float* r, float* y, const float K[27], float d,
                           const float* const m[6], const float* tot, int64_t D, int64_t H,
                           int64_t W, int fz, int fy, int fx, void* stream);

/* y = J(x - w) against b (same operator choice as above); flags
 * HFM_CORRECT_ZERO_OUTLET zeroes (x - w) on the outlet plane first
 * (UM/AI_HFM_v3.ipynb:2611-2613). */
int hfm_correct_jacobi_f32(const float* x, const float* w, const float* b, float* y,
                           const float K[27], float d, const float* const m[6], const float* tot,
                           int64_t D, int64_t H, int64_t W, int flags, void* stream);

/* y = J(x - P wc) against b, P = nearest repeat by (fz, fy, fx) of the coarse
 * correction wc (D/fz, H/fy, W/fx): the urban drivers' fine-level step, which
 * applies the prolonged correction with no fine-level sweep of it --
 * "w_512 = CNN3D_prol_256(w_256); values_p = values_p - w_512; outlet = 0;
 * values_p = values_p - CNN3D_A_512(values_p)/d + b/d" (UM/AI_HFM_v3.ipynb:2610-2613,
 * UM/INHALE_demo.ipynb:13454-13458, ARTs/naiver_stokes_3D.py:696-703).
 * Bitwise hfm_prolong_nearest_f32 followed by hfm_correct_jacobi_f32; flags
 * as there. */
int hfm_prolong_correct_jacobi_f32(const float* x, const float* wc, const float* b, float* y,
                                   const float K[27], float d, const float* const m[6],
                                   const float* tot, int64_t D, int64_t H, int64_t W, int fz,
                                   int fy, int fx, int flags, void* stream);

/* Explicit advection-diffusion step of nf (1 or 3) fields by the velocity
 * vel = (u, v, w):  out_q = f * (((dif(phi_q) - u ax(phi_q)) - v ay(phi_q)) -
 * w az(phi_q)) + base_q, with K = {dif, ax, ay, az} 27-tap kernels
 * (UM/AI_HFM_v3.ipynb cell 25: the momentum predictor / corrector with
 * CNN3D_central_2nd_dif/xadv/yadv/zadv, and the scalar transport).  Host
 * arrays of device pointers; outputs alias no input field. */
int hfm_advect27_f32(const float* const vel[3], const float* const* phi, const float* const* base,
                     float* const* out, int nf, const float K[4][27], float f, int64_t D,
                     int64_t H, int64_t W, void* stream);

/* hfm_advect27_f32 with the scalar's boundary fill and obstacle damping in
 * the same pass (the urban step's scalar transport, UM/AI_HFM_v3.ipynb cell 25:
 * advection, boundary_condition_indicator, then bluff_body before and after
 * the projection; alpha is not touched in between).  post: HFM_ADV_INDICATOR
 * (every boundary cell -> 0) | nbluff << HFM_ADV_BLUFF_SHIFT (0..2 dampings
 * out /= 1 + dt sigma).  Bitwise the separate hfm_advect27_f32, hfm_bc_f32
 * (HFM_BC_INDICATOR) and nbluff hfm_bluff_body_f32 calls. */
#define HFM_ADV_INDICATOR 1
#define HFM_ADV_BLUFF_SHIFT 4
int hfm_advect27_post_f32(const float* const vel[3], const float* const* phi,
                          const float* const* base, float* const* out, int nf,
                          const float K[4][27], float f, const float* sigma, float dt, int post,
                          int64_t D, int64_t H, int64_t W, void* stream);

/* As hfm_advect27_f32 with the diffusion stencil replaced by a given
 * dissipation field: out_q = f * (((dis_q - u ax(phi_q)) - v ay(phi_q)) -
 * w az(phi_q)) + base_q, K = (ax, ay, az) (UM/INHALE_demo.ipynb cell 24, the
 * momentum stages with the Petrov-Galerkin dissipation k_x, k_y, k_z). */
int hfm_advect27_dis_f32(const float* const vel[3], const float* const* phi,
                         const float* const* base, const float* const* dis, float* const* out,
                         int nf, const float K[3][27], float f, int64_t D, int64_t H, int64_t W,
                         void* stream);

/* Petrov-Galerkin dissipation (UM/INHALE_demo.ipynb cell 20,
 * Petrov_Galerkin_dissipation), on the damped velocities vt = f / (1 + dt
 * sigma) (hfm_bluff_body_f32).  Coefficients per component:
 *   k_q = min(8*0.25 |1/3 (|u~|+|v~|+|w~|) dif(vt_q)|
 *             / (eps + (|ax(vt_q)| + |ay(vt_q)| + |az(vt_q)|)/3), 0.5) / (1 + dt sigma)
 * with K = (dif, ax, ay, az); flux per component:
 *   out_q = 3*0.5 (k_q dif(vt_q) + dif(vt_q k_q) - vt_q dif(k_q)). */
int hfm_pg_coef_f32(const float* const vt[3], const float* sigma, float* const kout[3],
                    const float K[4][27], float dt, float eps, int64_t D, int64_t H, int64_t W,
                    void* stream);
int hfm_pg_flux_f32(const float* const vt[3], const float* const kin[3], float* const out[3],
                    const float Kdif[27], int64_t D, int64_t H, int64_t W, void* stream);

/* y = scale * sum_q K_q (*) x_q over nf <= 3 fields: the divergence source
 * b = -(CNN3D_Su(u) + CNN3D_Sv(v) + CNN3D_Sw(w)) (UM/AI_HFM_v3.ipynb:2568);
 * flags HFM_SUM_SOURCE_BC applies boundary_condition_source. */
int hfm_conv27_sum_f32(const float* const* x, float* y, int nf, const float K[3][27], float scale,
                       int64_t D, int64_t H, int64_t W, int flags, void* stream);

/* The urban step's pressure gradient, u_q -= K_q (*) p for the three velocity
 * components (UM/AI_HFM_v3.ipynb cell 25: values_u - CNN3D_Su(values_p), ...),
 * in one walk over p: o_q = f_q' - K_q (*) p' with
 *   HFM_GRAD_P_BC:       p' = boundary_condition_pressure(p), read through the
 *                        fill's cell map; pout (nullable, != p) receives p';
 *   HFM_GRAD_PRE_BLUFF:  f' = f / (1 + dt sigma) (bluff_body before the gradient);
 *   HFM_GRAD_POST_BLUFF: o_q /= 1 + dt sigma, except at the cells the velocity
 *                        fill copies from (u: j = 1, j = H-2, k = D-2; v: k = D-2),
 *                        for hfm_bc_velocity_bluff_inplace_f32 to finish.
 * o_q may be f_q (in place); the three kernels share a tap set.  Bitwise the
 * separate hfm_bc_f32, hfm_bluff_body_f32 and hfm_conv27_f32 (SUBTRACT) calls. */
#define HFM_GRAD_P_BC 1
#define HFM_GRAD_PRE_BLUFF 2
#define HFM_GRAD_POST_BLUFF 4
int hfm_grad3_f32(const float* p, const float* const f[3], float* const o[3], float* pout,
                  const float K[3][27], const float* sigma, float dt, int flags, int64_t D,
                  int64_t H, int64_t W, void* stream);

/* out = in with one boundary fill applied (kind = HFM_BC_*).  Copies read
 * the input field, as the reference's tf.Variable copies do; out != in. */
int hfm_bc_f32(const float* in, float* out, int64_t D, int64_t H, int64_t W, int kind,
               void* stream);

/* (uo, vo, wo) = velocity boundary fill of (u, v, w) (UM/AI_HFM_v3.ipynb
 * cell 19, boundary_condition_velocity); outputs must not alias inputs.
 * inflow (nullable, D values): u on the x faces per height instead of ub. */
int hfm_bc_velocity_f32(const float* u, const float* v, const float* w, float* uo, float* vo,
                        float* wo, int64_t D, int64_t H, int64_t W, float ub, const float* inflow,
                        void* stream);

/* The same fill in place: only the boundary planes are written (bitwise the
 * out-of-place result).  scratch: 2*D*W + 2*H*W floats of device memory (the
 * pre-fill rows and plane that boundary cells copy). */
int hfm_bc_velocity_inplace_f32(float* u, float* v, float* w, float* scratch, int64_t D, int64_t H,
                                int64_t W, float ub, const float* inflow, void* stream);

/* The in-place fill followed by bluff_body on the three fields (the end of the
 * urban step, UM/AI_HFM_v3.ipynb cell 25), after hfm_grad3_f32 with
 * HFM_GRAD_POST_BLUFF: boundary cells get the fill's value damped with their
 * own sigma, and the cells the fill copies from (left undamped by the
 * gradient) are damped from their gathered values.  Bitwise
 * hfm_bc_velocity_inplace_f32 then hfm_bluff_body_f32 after the plain gradient. */
int hfm_bc_velocity_bluff_inplace_f32(float* u, float* v, float* w, float* scratch, int64_t D,
                                      int64_t H, int64_t W, float ub, const float* inflow,
                                      const float* sigma, float dt, void* stream);

/* hfm_bc_velocity_inplace_f32 without a workspace, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  The pre-fill rows go to a scratch array
 * of HFM_BC_NOALLOC_SCRATCH_FLOATS floats (4 MiB: grids up to 512^3) that the
 * library holds on each device it runs on until the process ends; grids
 * needing more (2 D W + 2 H W > the maximum) return HFM_EINVAL.  It calls no
 * allocator.  Every call on one device shares that device's array: calls
 * that overlap on one GPU (different streams, or host threads) race.  The
 * stream must belong to the current device. */
#define HFM_BC_NOALLOC_SCRATCH_FLOATS (1 << 20)
int hfm_bc_velocity_inplace_noalloc_f32(float* u, float* v, float* w, int64_t D, int64_t H,
                                        int64_t W, float ub, const float* inflow, void* stream);

/* out_i = in_i / (1 + dt*sigma) for n fields (in place allowed).  Up to three
 * fields share a launch that reads sigma once per cell; when an output is
 * another field's input the fields run one launch each, in order. */
int hfm_bluff_body_f32(const float* const* in, float* const* out, int nfields,
                       const float* sigma, float dt, int64_t ncells, void* stream);

/* ---------------- config mode ---------------- */

/* One grid level of one instance, as seen by one launch.  The meaning of the
 * pointer slots depends on the kernel (see each entry point). */
typedef struct hfm_mg_task {
  int32_t D, H, W, inst;        /* extents of this level, instance index       */
  int32_t blk0, nblk, slot0;    /* first block, block count, first dot slot    */
  int32_t tile_rows;            /* fused pass pairs: rows per tile (hfm_mg_tile) */
  uint64_t magW, magH;          /* floor(2^64/W)+1, floor(2^64/H)+1 (0 if 1)   */
  int32_t Dc, Hc, Wc;           /* coarse extents (restriction / correction)   */
  int32_t tile_planes;          /* fused pass pairs: planes per tile            */
  uint64_t magWc, magHc;
  const void* c[6];             /* cx, cy, cz, bxh, byh, bzh                    */
  const void* xin;
  void* xout;
  const void* b;
  const void* ec;
  double* x64;
  double* r64;
  void* rlow;
  const float* b32;
  void* o[6];                   /* output faces of the batched setup kernels     */
} hfm_mg_task;

/* Per-instance solver state (device memory), FP64 flexible CG.  The host
 * stages a slot with active = 1 and iter = -1; HFM_FIN_INIT then applies to
 * staged slots only and the other finalize steps to initialised ones. */
typedef struct hfm_mg_scalars {
  double bnorm2, rz, alpha, beta;
  double rr, rel, tol, pap;
  int32_t iter, maxit, active, status; /* status: 0 running 1 converged 2 maxit 3 breakdown */
  int32_t slot0, nslot;
  int32_t xlag;       /* x lacks alpha_prev * p_prev: 0 no, 1 p_prev = the update's p (xin),
                         2 p_prev = the update's previous p (ec); see hfm_fcg_update */
  int32_t method;     /* HFM_METHOD_FCG or HFM_METHOD_REFINE */
  double alpha_prev;
  /* Residual scaling (FP16 storage): the fine right-hand side of the V-cycle
   * is T(2^sexp r); zexp is the exponent the current z = V(.) was made with
   * (the finalize divides r.z by 2^zexp: FCG is invariant under a rescaled
   * z, and refinement takes alpha = 2^-zexp).  With `scaled` set, INIT and
   * CHECK choose the next sexp = sbase - ilogb(||r||), so that the scaled
   * residual's RMS stays near 2^(sbase - log2(sqrt(N))).  All zero: no
   * scaling (FP32 / BF16 storage, bitwise the unscaled solver). */
  int32_t sexp, zexp, sbase, scaled;
} hfm_mg_scalars;

/* Outer iterations (hfm_mg_scalars.method).  REFINE is the FP64 defect
 * correction x += V(r), r -= A V(r) (iterative refinement with one V-cycle
 * per step): the same launches as FCG, with beta = 0 and alpha = 1 set by
 * hfm_fcg_finalize (the direction pass then copies z into p). */
#define HFM_METHOD_FCG 0
#define HFM_METHOD_REFINE 1

#define HFM_MG_RUNNING 0
#define HFM_MG_CONVERGED 1
#define HFM_MG_MAXIT 2
#define HFM_MG_BREAKDOWN 3

/* Sweep modes */
#define HFM_SWEEP_ZERO 0    /* out = omega b / d                              */
#define HFM_SWEEP_PLAIN 1   /* out = x + omega (b - A x)/d                    */
#define HFM_SWEEP_CORRECT 2 /* x := xin + alpha P ec, then a PLAIN sweep      */

/* Launch geometry shared by all work-table kernels: `blk_task[g]` is the
 * task index of block g (uint16), nblocks the grid size.  Tasks are padded to
 * a multiple of 8 blocks so that blocks of one task spread evenly over XCDs. */

/* Weighted-Jacobi sweep on the face operator (storage `store` for c, xin, b,
 * ec; `out_store` for xout: the fine level's last post-sweep may write the
 * preconditioned residual z wider than the cycle's internals).  Slots: c =
 * this level's coefficients, xin (unused for ZERO), b, xout, ec (coarse
 * correction for CORRECT, extents Dc/Hc/Wc).  `block`: threads per block
 * (128, 256 or 512; a block covers 8 * block cells of its task, the work
 * table's layout must use that chunk).  dots != NULL: partial sum(b*out) per
 * block into dots[slot0 + local block], out rounded to out_store; block must
 * then be 256 (the FCG kernels' chunk of 2048 cells). */
int hfm_mg_sweep(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                 const hfm_mg_scalars* sc, int store, int out_store, int mode, float omega,
                 float alpha, int block, double* dots, void* stream);

/* bc = R (b - A x): one coarse cell per thread over its <= 8 children,
 * R = P^T/8.  Slots: c, xin = x, b, xout = bc (coarse, extents Dc/Hc/Wc).
 * Blocks of `block` threads (128, 256, 512) cover 4 * block coarse cells. */
int hfm_mg_resid_restrict(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                          const hfm_mg_scalars* sc, int store, int block, void* stream);

/* Fused pass pairs of one level (one launch instead of two; the first
 * pass's result stays on chip).  Each task is cut into tiles of tile_rows
 * rows x tile_planes planes (fields of the task, from hfm_mg_tile); one
 * block per tile: task q owns the tiles blk0 .. blk0 + nblk - 1 (nblk = its
 * tile count), numbered contiguously over the tasks, and blk_task maps a
 * tile to its task, padded with 0xFFFF to nblocks (a multiple of 8).  Needs
 * W % 8 == 0.  512 threads per block.
 *   smooth_restrict: xout = the PLAIN sweep of xin against b, and
 *     rlow (coarse, extents Dc/Hc/Wc) = R (b - A xout): hfm_mg_sweep (PLAIN)
 *     then hfm_mg_resid_restrict, bitwise (the residual and its restrictions of
 *     ART/Anisotropic_3d_Multigrid_7point.ipynb:407-418, with a smoothing
 *     sweep before).
 *   smooth2: xout (type out_store) = PLAIN(CORRECT(xin, ec)) against b:
 *     hfm_mg_sweep (CORRECT) then hfm_mg_sweep (PLAIN), bitwise; with `dots`
 *     the b.xout partials of the instance's slots, one per tile (the others
 *     zeroed; the task's tile count must not exceed the instance's nslot).
 *     The prolongation + Jacobi steps of the same cycle (:420-443), with a
 *     second sweep after. */
int hfm_mg_smooth_restrict(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                           const hfm_mg_scalars* sc, int store, float omega, void* stream);
int hfm_mg_smooth2(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                   const hfm_mg_scalars* sc, int store, int out_store, float omega, float alpha,
                   double* dots, void* stream);

/* Tile geometry of the fused pass pairs for a level of extents (D, H, W):
 * rows and planes per tile and the tile count.  Tiles hold at most
 * max_planes planes (rounded to balanced, even counts; 0 = 32), as many rows
 * as a 512-thread block holds.  Returns HFM_EINVAL (with hfm_last_error) when
 * the level cannot be tiled (W % 8 != 0, or W too wide for a 512-thread block
 * to hold four rows). */
int hfm_mg_tile(int64_t D, int64_t H, int64_t W, int32_t max_planes, int32_t* tile_rows,
                int32_t* tile_planes, int64_t* ntiles);

/* Coarsest solve e = Ainv b (Ainv FP64 dense, N = D*H*W <= 64): one block
 * per task.  Slots: b, xout = e, ec = Ainv. */
int hfm_mg_coarse_solve(const hfm_mg_task* tasks, int64_t ntasks, const hfm_mg_scalars* sc,
                        int store, void* stream);

/* The V-cycle's tail in one launch: one 1024-thread workgroup per instance
 * runs V(nu1, nu2) from level lc (whose rhs B the caller's restriction wrote)
 * down to the coarsest (dense inverse) and back up, leaving level lc's
 * correction where the per-level launches would.  chain[2 g], chain[2 g + 1]:
 * the first record and the number of levels of instance g (device int32s);
 * record q: level lc + q with D/H/W, the next level's Dc/Hc/Wc and magics,
 * c = storage faces, xin = X, xout = Y, b = B, ec = the next level's B (the
 * coarsest: its FP64 inverse, N <= 64).  Per-cell arithmetic identical to
 * hfm_mg_sweep / hfm_mg_resid_restrict / hfm_mg_coarse_solve (their per-cell
 * path).  Replaces the per-level launches the sawtooth cycle of
 * ART/Anisotropic_3d_Multigrid_7point.ipynb:402-454 makes on small levels. */
int hfm_mg_coarse_cycle(const hfm_mg_task* recs, const int32_t* chain, int64_t ninst,
                        const hfm_mg_scalars* sc, int store, int nu1, int nu2, float omega,
                        float alpha, void* stream);

/* FCG kernels on the fine level (coefficients c in FP32).  T = `store` (the
 * V-cycle's storage: rlow and the fused first sweep), P = `pstore` (z and the
 * search direction p; hfm_fcg_dir's `store` argument is P).
 * init:   x64 = 0, r64 = b32, rlow = T(b32), dots = b.b  (staged slots only)
 * dir:    xout = p = P(b + beta xin) (b = z, xin = p_old), dots = p.(A p)
 * update: r64 -= alpha A p, rlow = T(r64), dots = r.r  (xin = p); the
 *         solution update is deferred to every other iteration: with xlag = 0
 *         x64 is left alone (HFM_FIN_CHECK then records alpha_prev = alpha,
 *         xlag = 1), with xlag = 1 x64 += alpha_prev p_prev, then += alpha p
 *         (p_prev in ec: the other search-direction buffer), in that order,
 *         so x64 is bitwise what updating every iteration gives.
 * flush:  for instances no longer active with xlag != 0, x64 += alpha_prev
 *         p_prev (xlag 1: p_prev in xin, 2: in ec); launched after every
 *         HFM_FIN_CHECK with the same update tasks (the next HFM_FIN_BETA
 *         clears xlag of inactive slots).
 * *_sweep variants also write the next V-cycle's first fine-level sweep from
 * a zero guess, xout = omega rlow / d (d = diagonal of the faces rounded to
 * storage), so that sweep is not launched separately. */
int hfm_fcg_init(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                 const hfm_mg_scalars* sc, int store, double* dots, void* stream);
int hfm_fcg_dir(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                const hfm_mg_scalars* sc, int store, double* dots, void* stream);
int hfm_fcg_update(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                   const hfm_mg_scalars* sc, int store, int pstore, double* dots, void* stream);
int hfm_fcg_init_sweep(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                       const hfm_mg_scalars* sc, int store, float omega, double* dots,
                       void* stream);
int hfm_fcg_update_sweep(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                         const hfm_mg_scalars* sc, int store, int pstore, float omega,
                         double* dots, void* stream);
int hfm_fcg_flush(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                  const hfm_mg_scalars* sc, int pstore, void* stream);

/* Per-instance fixed-order reduction of dots[slot0 .. slot0+nslot) and the
 * scalar step:  kind 0 INIT (bnorm2), 1 BETA (rz, beta), 2 ALPHA (pap,
 * alpha), 3 CHECK (rr, rel, iter, status).  One block per instance.  With
 * method REFINE, BETA sets beta = 0 and ALPHA alpha = 1 (pap is only checked
 * to be finite). */
#define HFM_FIN_INIT 0
#define HFM_FIN_BETA 1
#define HFM_FIN_ALPHA 2
#define HFM_FIN_CHECK 3
int hfm_fcg_finalize(hfm_mg_scalars* sc, int64_t ninst, const double* dots, int kind,
                     void* stream);

/* Setup kernels.
 * faces:    cx/cy/cz (D,H,W) and bxh (D,H), byh (D,W), bzh (H,W) from cell
 *           conductivities sx, sy, sz (harmonic face means, boundary face =
 *           cell value), all FP32, scaled by inv_h2.
 * galerkin: coarse faces (ceil halves) = sum of crossing fine faces / 8.
 * convert:  n FP32 values -> storage `store`.
 * coarse_inverse: dense FP64 inverse of the face operator of each task's
 *           level (N <= 64) into the task's xout (double[N*N]); c = FP32 faces. */
int hfm_mg_faces(const float* sx, const float* sy, const float* sz, float* const out[6],
                 int64_t D, int64_t H, int64_t W, float inv_h2, void* stream);
int hfm_mg_galerkin(const float* const fine[6], float* const coarse[6], int64_t D, int64_t H,
                    int64_t W, void* stream);
int hfm_mg_convert(const float* src, void* dst, int64_t n, int store, void* stream);

/* Batched setup over a work table (blk_task as for the cell kernels, tasks
 * need not be padded to 8 blocks):
 * faces_batched:    per task xin = sx, b = sy, ec = sz (FP32 cell
 *                   conductivities of extents D/H/W), o[0..5] = output faces;
 *                   a block covers 2048 cells (8 per thread; 16-byte rows when
 *                   W % 8 == 0, so the arrays must be 16-byte aligned then).
 * galerkin_batched: per task D/H/W = fine extents, c[0..5] = fine FP32 faces,
 *                   o[0..5] = coarse FP32 faces; 256 coarse cells per block. */
int hfm_mg_faces_batched(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                         float inv_h2, void* stream);
int hfm_mg_galerkin_batched(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                            void* stream);
/* convert_batched:  per task xin = FP32 source, xout = destination of storage
 *                   `store`, D = element count; a block covers 2048 elements. */
int hfm_mg_convert_batched(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                           int store, void* stream);
int hfm_mg_coarse_inverse(const hfm_mg_task* tasks, int64_t ntasks, void* stream);

/* ---------------- persistent iteration ----------------
 * hfm_mg_persistent runs a program of phases in one cooperative launch: a
 * phase is one of the launches above (its table, block map and arguments;
 * the cell kernels with 256-thread blocks), its work items that launch's
 * blocks (for HFM_PH_FINALIZE the scalar slots slot0 .. slot0 + nitems - 1,
 * for HFM_PH_COARSE_CYCLE the chain's instances).  For each instance slot the
 * phases run in program order: an item of phase p for slot s starts once all
 * need[q][s] items of phase q = prev[p][s] for s have finished (prev -1: no
 * wait), so the results are bitwise those of the launches run one after the
 * other.  Arrays are [nphase][nslot] in device memory; cnt must be zero at the
 * launch.  A wait longer than timeout_s (wall clock) sets *fail and ends the
 * launch early (results void).  workgroups: a multiple of 8, at most
 * hfm_mg_persistent_workgroups (every workgroup must be resident).  No
 * reference counterpart: the reference's multigrid cycles are host loops of
 * layer calls (ART/Anisotropic_3d_Multigrid_7point.ipynb:402-454). */
#define HFM_PH_SWEEP 0           /* hfm_mg_sweep (mode, dot; out = pstore if dot)   */
#define HFM_PH_RESTRICT 1        /* hfm_mg_resid_restrict                          */
#define HFM_PH_SMOOTH_RESTRICT 2 /* hfm_mg_smooth_restrict                         */
#define HFM_PH_SMOOTH2 3         /* hfm_mg_smooth2 (dot; out = pstore if dot)       */
#define HFM_PH_COARSE_CYCLE 4    /* hfm_mg_coarse_cycle (tasks = records, chain)    */
#define HFM_PH_FINALIZE 5        /* hfm_fcg_finalize (mode = kind)                  */
#define HFM_PH_DIR 6             /* hfm_fcg_dir                                     */
#define HFM_PH_UPDATE 7          /* hfm_fcg_update_sweep                            */
#define HFM_PH_FLUSH 8           /* hfm_fcg_flush                                   */

typedef struct hfm_phase {
  int32_t kind, mode, dot, nitems;
  int32_t nu1, nu2, slot0, pad;
  float omega, alpha;
  const hfm_mg_task* tasks;
  const uint16_t* blk_task;
  const int32_t* chain;
  double* dots;
} hfm_phase; /* 72 bytes */

typedef struct hfm_persist_ctl {
  uint32_t* cnt;
  const int32_t* need;
  const int32_t* prev;
  uint32_t* fail;
  int32_t nslot;
  int32_t pad;
} hfm_persist_ctl;

int hfm_mg_persistent(const hfm_phase* prog, int32_t nphase, const hfm_persist_ctl* ctl,
                      hfm_mg_scalars* sc, int store, int pstore, int32_t workgroups,
                      double timeout_s, void* stream);
/* Workgroups of hfm_mg_persistent that are resident at once on the current
 * device (a multiple of 8). */
int hfm_mg_persistent_workgroups(int store, int pstore, int32_t* n);
/* sizeof(hfm_phase) and sizeof(hfm_persist_ctl) as compiled. */
int hfm_mg_persist_struct_sizes(int64_t* phase_size, int64_t* ctl_size);

/* ---------------- 2D periodic flow (Kolmogorov configuration) ----------------
 * The reference's 2D layers (explicit two-stage stencil update,
 * ARTn/Advection_diffusion.ipynb:291-296; mean restriction / nearest
 * prolongation / Jacobi multigrid, ARTn/Parallel_multigrid_2D.ipynb#c5-#c20)
 * on an N x N periodic staggered grid: u on x-faces, v on y-faces, phi at
 * cell centres, arrays row-major [j][i].  One workgroup (1024 threads) runs
 * a whole step of one flow; a launch advances `nflows` independent flows.
 *
 * Per step of dt: n = max(1, ceil((max|u| + max|v|) * cfl_coef)) substeps of
 * dt/n (cfl_coef = dt / (h cfl_max)), each
 *   uh = u + dt/2 F(u);  u* = u + dt F(uh)      (F: 5-point diffusion,
 *        velocity-signed first-order upwind advection, forcing fu[j] on u)
 *   b  = h (u*_i - u*_{i+1} + v*_j - v*_{j+1}) - mean(b)
 *   A phi = b, A = 4 phi - sum of the 4 neighbours, by FP64 CG (phi, r FP64,
 *        warm start) preconditioned by one FP32 V(nu1, nu2) cycle: weighted
 *        Jacobi, 2x2 mean restriction, nearest prolongation times alpha,
 *        level l operator 2^-l A, coarsest level by the FP64 matrix `pinv`
 *        (the pseudo-inverse of the coarsest level's operator, n_c^2 x n_c^2,
 *        n_c = N >> (nlev-1), n_c^2 <= 1024);
 *   u = u* - (phi - phi_w)/h,  v = v* - (phi - phi_s)/h.
 * Stopping rule per substep: ||r||_2 <= tol ||b||_2.  A non-finite velocity,
 * right-hand side or residual stops the flow with HFM_MG_BREAKDOWN.
 *
 * Buffers per flow (device): u, v, uh, vh, z, p: N*N floats; phi, r: N*N
 * doubles; mgx, mgy, mgb: sum over levels of (N >> l)^2 floats; fu: N floats;
 * stats: HFM_FLOW_STATS doubles. */
#define HFM_FLOW_STATS 8
#define HFM_FLOW_ST_ITERS 0      /* CG iterations of the last substep         */
#define HFM_FLOW_ST_REL 1        /* its final relative residual               */
#define HFM_FLOW_ST_ITERS_SUM 2  /* iterations summed over all substeps        */
#define HFM_FLOW_ST_STATUS 3     /* HFM_MG_* of the last substep; HFM_FLOW_CFL */
#define HFM_FLOW_ST_STEPS 4      /* steps completed                           */
#define HFM_FLOW_ST_SUBSTEPS 5   /* substeps summed                           */
#define HFM_FLOW_ST_REL_MAX 6    /* largest final relative residual            */
#define HFM_FLOW_ST_NSUB 7       /* substeps of the last step                  */
#define HFM_FLOW_CFL 4           /* status: more than max_sub substeps needed */
#define HFM_FLOW_TEAM_TIMEOUT 5  /* status: a team barrier waited more than 1 s of
                                    wall-clock time (fields void) */
#define HFM_FLOW_TEAM_BYTES 4096 /* team scratch per flow */
#define HFM_FLOW_PROJECT_ONLY 1  /* flag: skip the momentum stages (project u, v) */

typedef struct hfm_flow2d_task {
  float* u;
  float* v;
  float* uh;
  float* vh;
  float* z;
  float* p;
  double* phi;
  double* r;
  float* mgx;
  float* mgy;
  float* mgb;
  const float* fu;
  double* stats;
  void* team;     /* team scratch, HFM_FLOW_TEAM_BYTES zeroed before each
                     launch; used when params.team > 1 */
} hfm_flow2d_task; /* 112 bytes */

typedef struct hfm_flow2d_params {
  float nu_h2;    /* nu / h^2 */
  float inv_h;    /* 1 / h */
  float h;
  float dt;
  float cfl_coef; /* dt / (h cfl_max) */
  float omega;
  float alpha;
  int32_t N;      /* grid edge, a power of two; the same for every flow */
  int32_t nu1;
  int32_t nu2;
  int32_t nlev;   /* levels N, N/2, ..., N >> (nlev-1) >= 2 */
  int32_t maxit;
  int32_t max_sub;
  int32_t team;   /* workgroups per flow, 1..64 (> 1: one cooperative launch) */
  double tol;
} hfm_flow2d_params; /* 64 bytes */

/* Advance every flow of `tasks` (device array) by `nsteps` steps (flags:
 * HFM_FLOW_PROJECT_ONLY projects u, v once instead).  A flow whose status is
 * not converged stops stepping.  params.team workgroups share each flow
 * (rank = block % team); with team > 1 the launch is cooperative (all
 * nflows * team workgroups co-resident, else HFM_EDEVICE) and every task's
 * team scratch must be zeroed before the call.  No reference counterpart: the reference's
 * time-step drivers are Python loops over Keras layer calls
 * (UM/AI_HFM_v3.ipynb cell 25). */
int hfm_flow2d_steps(const hfm_flow2d_task* tasks, int64_t nflows, const hfm_flow2d_params* prm,
                     const double* pinv, int nsteps, int flags, void* stream);
/* sizeof(hfm_flow2d_task) and sizeof(hfm_flow2d_params) as compiled. */
int hfm_flow2d_struct_sizes(int64_t* task_size, int64_t* params_size);

#ifdef __cplusplus
}
#endif
#endif /* HFM_H */
