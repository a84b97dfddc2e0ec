/*
 * tlbm_hip.h — C ABI of libtlbm_hip.so, the MI355X (gfx950) per-step lattice
 * update of TensorLBM.
 *
 * Conventions shared by every entry point:
 *   - Populations are SoA ``f[Q][nz][ny][nx]``, x fastest, contiguous (the
 *     reference layout, src/tensorlbm/d3q19.py:95-103).  D2Q9 fields
 *     ``f[9][ny][nx]`` are passed with nz = 1.
 *   - All pointers are device pointers owned by the caller; the library
 *     allocates no device memory.  Its one device array of its own is the
 *     2.76 MB static workspace of tlbm_mlp_train_step_noalloc.  Work is
 *     enqueued on ``stream`` (a hipStream_t; 0 is the null stream) and nothing
 *     synchronises.
 *   - Streaming entry points require ``f_in != f_out``: a pull step reads
 *     neighbouring cells, so aliased buffers race across workgroups
 *     (reference src/tensorlbm/triton_fused.py:697-701).
 *   - Return value: TLBM_OK or a negative TLBM_ERR_* code;
 *     tlbm_last_error() returns a message for the calling thread.
 */
#ifndef TLBM_HIP_H
#define TLBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tlbm_stream_t; /* == hipStream_t */

/* status codes */
#define TLBM_OK 0
#define TLBM_ERR_INVALID (-1)     /* bad argument value */
#define TLBM_ERR_SHAPE (-2)       /* inconsistent or oversized shape */
#define TLBM_ERR_DTYPE (-3)       /* unsupported storage / dtype */
#define TLBM_ERR_UNSUPPORTED (-4) /* lattice x collision combination not built */
#define TLBM_ERR_ALIAS (-5)       /* f_in == f_out for a streaming op */
#define TLBM_ERR_HIP (-100)       /* HIP runtime error (launch, symbol, ...) */

/* lattices (value = Q) */
#define TLBM_D2Q9 9
#define TLBM_D3Q19 19
#define TLBM_D3Q27 27

/* population storage */
#define TLBM_STORE_F32 0
#define TLBM_STORE_F64 1
#define TLBM_STORE_F16 2      /* plain f in binary16, fp32 arithmetic */
#define TLBM_STORE_BF16 3     /* plain f in bfloat16, fp32 arithmetic */
#define TLBM_STORE_F16_DEV 4  /* f_i - w_i in binary16, fp32 arithmetic */
#define TLBM_STORE_BF16_DEV 5 /* f_i - w_i in bfloat16, fp32 arithmetic */

/* collision families */
#define TLBM_COLL_NONE 0     /* stream only */
#define TLBM_COLL_BGK 1      /* solver.py:67-71, solver3d.py:100-104, d3q27.py:423-435 */
#define TLBM_COLL_MRT 2      /* solver.py:249-335, solver3d.py:107-160, d3q27.py:643-726 */
#define TLBM_COLL_CM 3       /* cascaded central moments, cascaded_collision.py:445-509;
                                D2Q9: collide_cumulant_d2q9, cumulant.py:116-311 */
#define TLBM_COLL_CUMULANT 4 /* D3Q19 Hermite-regularised, cumulant.py:1159-1377;
                                D3Q27 legacy operator on f_neq, cumulant.py:935-1047 */
#define TLBM_COLL_GEIER 5    /* D3Q27 Geier 2015 cumulants, cumulant.py:1050-1156 */

/* source of the shear relaxation rate */
#define TLBM_TAU_CONST 0 /* omega = 1/tau */
#define TLBM_TAU_SMAG 1  /* clamped Hou closure, turbulence.py:183-208 */
#define TLBM_TAU_FIELD 2 /* per-cell tau_eff array, triton_fused_obstacle.py:1300-1312 */
#define TLBM_TAU_SMAG_RAW 3 /* unclamped Hou closure of cumulant.py:1316-1321 */
#define TLBM_TAU_MLP 4   /* learned closure: tau_eff = max(tau + 3 nu_t, 0.501) (ai/inference.py:56-65),
                            nu_t from the tlbm_mlp_pack blob in coll->mlp (D3Q19 BGK) */
#define TLBM_TAU_FIELD64 5 /* per-cell tau_eff array of float64 values (coll->tau_eff points to doubles);
                              tlbm_collide only -- the fp64 form of TLBM_TAU_FIELD */
#define TLBM_TAU_MLP_BF16 6 /* TLBM_TAU_MLP with the bf16-operand network: coll->mlp from tlbm_mlp_pack_bf16
                               (D3Q19 BGK; fp32, fp64, fp16- or bf16-deviation storage) */

/* gradient eddy-viscosity closures for tlbm_gradient_closure_tau */
#define TLBM_SGS_WALE 1   /* _wale_nu_t_3d, turbulence.py:754-841 */
#define TLBM_SGS_VREMAN 2 /* _vreman_nu_t_3d, turbulence.py:844-872 */

/* boundary kinds for tlbm_step */
#define TLBM_BC_NONE 0
#define TLBM_BC_CHANNEL_X 1 /* equilibrium inlet x=0 (rho from x=1), copy outlet x=nx-1,
                               boundaries3d.py:407-443 */
#define TLBM_BC_LID_Y 2     /* Zou-He moving lid on the fluid cells of y=ny-1 (solid cells there
                               keep bounce-back), boundaries3d.py:324-404 */
#define TLBM_BC_FAR_FIELD 3 /* free-stream equilibrium (rho = 1, u) on the inlet x=0 and the lateral
                               faces named in `faces`, copy outlet x=nx-1 (far_field_bc_3d,
                               boundaries3d.py:446-537; lateral faces win on shared edges) */

/* tlbm_boundary.faces bits for TLBM_BC_FAR_FIELD */
#define TLBM_FACE_YM 1           /* y = 0 plane */
#define TLBM_FACE_YP 2           /* y = ny-1 */
#define TLBM_FACE_ZM 4           /* z = 0 */
#define TLBM_FACE_ZP 8           /* z = nz-1 */
#define TLBM_FACE_XM_PERIODIC 16 /* leave the inlet plane to the periodic stream */
#define TLBM_FACE_XP_PERIODIC 32 /* leave the outlet plane to the periodic stream */

typedef struct tlbm_collision {
  int32_t kind;         /* TLBM_COLL_* */
  int32_t tau_mode;     /* TLBM_TAU_* */
  int32_t eq_rho_floor; /* 1: f_eq built with max(rho,1e-12) (fused kernels); 0: with rho */
  int32_t reserved;
  double tau;           /* molecular relaxation time */
  double cs_delta_sq;   /* Cs^2 Delta^2 for the Smagorinsky modes; TLBM_TAU_MLP: the strain
                           features use the Smagorinsky tau of this Cs as time scale
                           (0: the molecular tau, _strain_from_fneq_3d, turbulence.py:1449-1480) */
  double rate[8];       /* MRT: s_e, s_eps, s_q, s_pi | CM: s_bulk, s_3, s_4, s_5, s_6 |
                           CUMULANT/GEIER: omega_b, omega_odd, omega_even */
  const float* tau_eff; /* TLBM_TAU_FIELD: (nz,ny,nx) float32; TLBM_TAU_FIELD64: float64 (cast) */
  const void* mlp;      /* TLBM_TAU_MLP: device pointer to packed MLP parameters */
} tlbm_collision;

typedef struct tlbm_boundary {
  int32_t kind;   /* TLBM_BC_* */
  int32_t faces;  /* TLBM_BC_FAR_FIELD: TLBM_FACE_* bits; 0 otherwise */
  double u[3];    /* inlet / lid / free-stream velocity */
} tlbm_boundary;

/* library */
int tlbm_version(void);
const char* tlbm_last_error(void);
const char* tlbm_status_string(int status);

/*
 * Population layouts: ``strides`` point to 3 int64 element strides (q, z, y);
 * the x stride is 1.  NULL means the reference's contiguous f[Q][nz][ny][nx].
 * One z-plane (y*sy + x) must span less than 4 GiB.  Input and output of a
 * streaming op may use different layouts (e.g. the solver's padded
 * [z][q][y][x] buffers); they must not overlap.
 */

/*
 * Fused pull-stream + boundary + collide, one launch
 *   (replaces triton_fused()          src/tensorlbm/triton_fused.py:618-667,
 *             triton_fused_obstacle_xfar_les()  triton_fused_obstacle.py:1695-2053).
 * solid  : optional int8 (nz,ny,nx) contiguous mask, non-zero = solid; full-way bounce-back.
 * bc     : optional boundary (NULL = periodic).
 * force  : optional 3 doubles on the device; the Ladd force of the post-stream,
 *          pre-bounce-back populations on solid cells is ADDED (caller zeroes).
 */
int tlbm_step(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
              const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll,
              const int8_t* solid, const tlbm_boundary* bc, double* force, tlbm_stream_t stream);

/*
 * Two consecutive tlbm_step()s in one launch, f_in -> f_out (temporal blocking: the
 * intermediate lattice stays in LDS).  Replaces two iterations of the reporter-free
 * run() loop, src/tensorlbm/triton_fused.py:892-895.  Same arguments and results as
 * tlbm_step(f_in -> tmp) then tlbm_step(tmp -> f_out), up to rounding.  D3Q19 /
 * D3Q27; F32, F16_DEV, BF16_DEV storage; every collision and boundary except
 * TAU_MLP.  force (D3Q19 only, needs solid): the Ladd force of BOTH steps is ADDED,
 * the sum of what the two tlbm_step() launches add (the per-step force phase of
 * triton_fused_obstacle.py:1494-1512, run every step).  Returns
 * TLBM_ERR_UNSUPPORTED for other arguments, which tlbm_step2_supported()
 * reports (force aside) without launching.
 */
int tlbm_step2(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
               const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll,
               const int8_t* solid, const tlbm_boundary* bc, double* force, tlbm_stream_t stream);

/* 1 if tlbm_step2() serves these arguments, else 0; sets no error */
int tlbm_step2_supported(int lattice, int store, const tlbm_collision* coll, const tlbm_boundary* bc,
                         int nx);

/* periodic pull stream (solver3d.py:210-239 stream3d, solver.py:396-417, d3q27.py:766-791) */
int tlbm_stream(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
                const int64_t* out_strides, int nz, int ny, int nx, tlbm_stream_t stream);

/* local collision, f_out may equal f_in (collide_bgk3d, collide_mrt3d, ...); F32 / F64 */
int tlbm_collide(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
                 const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll,
                 tlbm_stream_t stream);

/* equilibrium (d2q9.py:39-58, d3q19.py:83-125, d3q27.py:353-392); dtype F32 or F64;
 * contiguous fields; uz may be NULL (D2Q9). */
int tlbm_equilibrium(int lattice, int dtype, const void* rho, const void* ux, const void* uy,
                     const void* uz, void* f_out, int64_t ncells, tlbm_stream_t stream);

/* macroscopic moments (d2q9.py:61-74, d3q19.py:165-186, d3q27.py:395-420); contiguous;
 * uz may be NULL. */
int tlbm_macroscopic(int lattice, int dtype, const void* f, void* rho, void* ux, void* uy, void* uz,
                     int64_t ncells, tlbm_stream_t stream);

/* full-way swap where(mask, f[opp], f) (boundaries3d.py:112-132); contiguous;
 * f_out may equal f_in */
int tlbm_bounce_back(int lattice, int dtype, const void* f_in, void* f_out, const int8_t* mask,
                     int64_t ncells, tlbm_stream_t stream);

/* ADD to out[0..3] (device doubles): sum f, sum c_x f, sum c_y f, sum c_z f (FP64 wave
 * reductions, correct_mass3d's sum, solver3d.py:287-300).  Any storage and layout;
 * deviation storages add back w. */
int tlbm_moment_sums(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny,
                     int nx, double* out, tlbm_stream_t stream);

/* ADD the Ladd force 2 sum_solid sum_q c_q f to out[0..3] (obstacles.py:91-130), FP64. */
int tlbm_obstacle_force(int lattice, int store, const void* f, const int64_t* strides, int nz,
                        int ny, int nx, const int8_t* solid, double* out, tlbm_stream_t stream);

/* ADD sum over cells of the clamped Smagorinsky tau_eff of each cell's non-equilibrium stress
 * (turbulence.py:183-208 with |Pi| of :143-180 / :436-469) to out[0] (device double, FP64).
 * collide_cumulant_d3q27's C_s > 0 path relaxes with the domain mean (cumulant.py:986-997). */
int tlbm_smagorinsky_tau_sum(int lattice, int store, const void* f, const int64_t* strides, int nz,
                             int ny, int nx, double tau, double cs_delta_sq, double* out,
                             tlbm_stream_t stream);

/*
 * Per-cell tau_eff of a velocity-gradient closure (the C_w / C_v branches of
 * collide_cumulant_d3q19, cumulant.py:1263-1303): the macroscopic velocity (ux, uy, uz: contiguous
 * (nz, ny, nx), float32 or float64 by dtype) takes wall_velocity[3] (host doubles; NULL = 0) on cells
 * where the optional int8 mask `solid` is non-zero, is differentiated with centred differences inside
 * and second-order one-sided differences on the faces (_velocity_gradients_3d, turbulence.py:595-695),
 * and tau_out = clip(tau + 3 nu_t, 0.5001, max(1, tau + 0.5)) (_nu_t_to_tau_eff, turbulence.py:875-888)
 * with nu_t of model TLBM_SGS_WALE (coef = C_w) or TLBM_SGS_VREMAN (coef = C_v); tau_out has dtype.
 * Feed it to tlbm_collide / tlbm_step as TLBM_TAU_FIELD (float32) or TLBM_TAU_FIELD64.
 */
int tlbm_gradient_closure_tau(int dtype, const void* ux, const void* uy, const void* uz, const int8_t* solid,
                              const double* wall_velocity, int nz, int ny, int nx, int model, double coef, double tau,
                              void* tau_out, tlbm_stream_t stream);

/*
 * Boundary conditions of a post-stream D3Q19 state as a separate op, out of place (f_out != f_in,
 * contiguous (19, nz, ny, nx), float32 or float64 by dtype).  Every value read is the input's, and
 * for each cell:
 *   bc->kind TLBM_BC_CHANNEL_X: x = 0 becomes f_eq(rho of x = 1, (u[0], 0, 0)), x = nx-1 the state
 *            of x = nx-2 (apply_simple_channel_boundaries_3d, boundaries3d.py:407-443);
 *   bc->kind TLBM_BC_FAR_FIELD: f_eq(1, u) on the inlet (unless TLBM_FACE_XM_PERIODIC) and on the
 *            lateral faces named in bc->faces, which win on shared edges; the outlet copies
 *            x = nx-2 (unless TLBM_FACE_XP_PERIODIC) (far_field_bc_3d, boundaries3d.py:446-537);
 *   bc->kind TLBM_BC_LID_Y: Zou-He with u_x = u[0] on every cell of y = ny-1, corners included
 *            (zou_he_moving_lid_3d, boundaries3d.py:324-404);
 *   bc NULL or TLBM_BC_NONE: no face condition;
 * then full-way bounce-back where the optional int8 mask_a is non-zero, then where mask_b is
 * (bounce_back_cells_3d applied twice: a cell in both masks swaps back).  Inlet / outlet kinds
 * need nx >= 3.
 */
int tlbm_boundary_eager(int dtype, const void* f_in, void* f_out, int nz, int ny, int nx,
                        const tlbm_boundary* bc, const int8_t* mask_a, const int8_t* mask_b,
                        tlbm_stream_t stream);

/* sphere_mask (boundaries3d.py:29-49): out[z][y][x] = 1 where (x-cx)^2 + (y-cy)^2 + (z-cz)^2 <=
 * radius^2, else 0, evaluated in float32 as the reference does (centre rounded to float32, squares
 * and sums rounded in turn, radius^2 formed in double and rounded to float32); out holds nz*ny*nx
 * bytes (a torch.bool tensor's storage). */
int tlbm_sphere_mask(int nx, int ny, int nz, double cx, double cy, double cz, double radius,
                     uint8_t* out, tlbm_stream_t stream);

/* f <- f * (*scale) (correct_mass3d, solver3d.py:287-300); scale is a device double. */
int tlbm_scale(int lattice, int store, void* f, const int64_t* strides, int nz, int ny, int nx,
               const double* scale, tlbm_stream_t stream);

/*
 * Learned eddy-viscosity closure, 3-D 7 -> 32 -> 32 -> 1 MLP (ReLU, Softplus(beta=10) output).
 * The reference's closure is 2-D only (EddyViscosityMLP, src/tensorlbm/ai/model.py:19-92,
 * trained by train_eddy_viscosity_model, ai/train.py:310-464, applied by collide_ai_les_bgk,
 * ai/inference.py:68-92); this 3-D form has no counterpart there.  Parameters are one fp32
 * vector of ntotal floats: W1[32][7] b1[32] W2[32][32] b2[32] w3[32] b3 | mean[7] inv_std[7]
 * (the first nparam are trainable, the last 14 normalise the inputs).
 * Features: S_xx S_yy S_zz S_xy S_xz S_yz |S| of the resolved strain rate; output nu_t >= 0.
 */
/* sizes of the parameter vector, the packed matrix-core blob and the gradient block */
int tlbm_mlp_sizes(int64_t* nparam, int64_t* ntotal, int64_t* packed_bytes, int64_t* batch_block);

/* fp32 parameters -> packed fp16 matrix-core blob used by tlbm_step with TLBM_TAU_MLP (coll->mlp) */
int tlbm_mlp_pack(const float* params, void* packed, tlbm_stream_t stream);

/* the same blob with bf16 operands ("BF16 weights, FP32 accumulation"): W1 and W2 as bf16 values,
 * the inputs, the layer-1 bias and the layer-1 activations as two-term bf16 splits; used by
 * tlbm_step with TLBM_TAU_MLP_BF16.  Same size as the fp16 blob. */
int tlbm_mlp_pack_bf16(const float* params, void* packed, tlbm_stream_t stream);

/* y[n] = MLP(x[n][7]) in fp32 (predict_tau_eff_2d's network call, ai/inference.py:53-65) */
int tlbm_mlp_predict(const float* params, const float* x, int64_t n, float* y, tlbm_stream_t stream);

/* MSE gradient partials: partials[ceil(n / batch_block)][nparam + 1] (last column: loss share) */
int tlbm_mlp_grad(const float* params, const float* x, const float* t, int64_t n, float* partials,
                  tlbm_stream_t stream);

/* sum the partials in a fixed order, clip the gradient's global 2-norm to clip_norm
 * (<= 0: no clipping; clip_grad_norm_ in ai/train.py:401-405) and apply one Adam step
 * (torch.optim.Adam rule, ai/train.py:382-435).  lr = 0 only reduces.  loss_out (MSE) and
 * grad_out (the unclipped gradient, nparam floats) may be NULL. */
int tlbm_mlp_adam(float* params, const float* partials, int nblocks, float* m, float* v, int step,
                  float lr, float beta1, float beta2, float eps, float clip_norm, float* loss_out,
                  float* grad_out, tlbm_stream_t stream);

/*
 * tlbm_mlp_grad + tlbm_mlp_adam in one call that allocates no device memory: same arguments as
 * those two (nblocks is derived from n), same kernels, same results bit for bit.  No counterpart
 * in the reference; it is for callers that cannot allocate (code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc).  The gradient partials live in a library-owned
 * device array of TLBM_MLP_NOALLOC_MAX_SAMPLES / batch_block blocks (2.76 MB, one copy per GPU,
 * held until the process ends); n above TLBM_MLP_NOALLOC_MAX_SAMPLES returns TLBM_ERR_SHAPE
 * without launching.  Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (several streams or host threads) race and must be serialised by the caller.
 */
#define TLBM_MLP_NOALLOC_MAX_SAMPLES 32768
int tlbm_mlp_train_step_noalloc(float* params, const float* x, const float* t, int64_t n, float* m,
                                float* v, int step, float lr, float beta1, float beta2, float eps,
                                float clip_norm, float* loss_out, float* grad_out,
                                tlbm_stream_t stream);

/*
 * The reference's 2-D learned closure (EddyViscosityMLP, ai/model.py:19-92): an MLP in -> H -> ... -> H -> 1
 * with hidden activation tanh / relu / gelu (exact erf form) and a Softplus(beta = 10) output.  Parameters are
 * one fp32 vector in nn.Sequential order -- per hidden layer W [H][in] then b [H], then the output w [H] and
 * b [1] -- followed by the input normalisation mean [in] and std [in] (inputs (x - mean) / std).
 * Built for hidden_features 8, 16, 32, 64; hidden_layers 1..4; in_features 1..8 (3 for the strain features).
 */
#define TLBM_ACT_TANH 0
#define TLBM_ACT_RELU 1
#define TLBM_ACT_GELU 2
typedef struct tlbm_mlp2d_arch {
  int32_t in_features, hidden_features, hidden_layers, activation;
} tlbm_mlp2d_arch;

/* number of trainable parameters (the normalisation's 2 * in_features follow them); -1 if not built */
int64_t tlbm_mlp2d_nparam(const tlbm_mlp2d_arch* arch);

/* y[n] = MLP(x[n][in]) in fp32 (EddyViscosityMLP.forward, ai/model.py:83-92) */
int tlbm_mlp2d_forward(const tlbm_mlp2d_arch* arch, const float* params, const float* x, int64_t n,
                       float* y, tlbm_stream_t stream);

/* periodic central-difference strain of a (ny, nx) velocity field (strain_rate_tensor_2d,
 * ai/dataset.py:41-72): feats[ny*nx][3] = S_xx, S_yy, S_xy; target (optional) = c_s^2 |S|,
 * |S| = sqrt(2 S:S) (extract_les_samples_2d, ai/dataset.py:89-127) */
int tlbm_strain2d(const float* ux, const float* uy, int ny, int nx, float* feats, float* target,
                  double c_s, tlbm_stream_t stream);

/* per cell of the velocity field: mode 0 nu_t = max(MLP(S), 0) (predict_nu_t_2d); mode 1
 * tau_eff = max(tau + 3 nu_t, 0.501) (predict_tau_eff_2d, ai/inference.py:53-65); out (ny, nx) */
int tlbm_mlp2d_nu_t(const tlbm_mlp2d_arch* arch, const float* params, const float* ux, const float* uy,
                    int ny, int nx, int mode, double tau, float* out, tlbm_stream_t stream);

/* collide_ai_les_bgk (ai/inference.py:68-92): f_out = f - (f - f_eq(rho, u)) / tau_eff with tau_eff of
 * each cell's strain; rho, ux, uy are tlbm_macroscopic of f (D2Q9, float32, contiguous (9, ny, nx)) */
int tlbm_ai_les_bgk(const tlbm_mlp2d_arch* arch, const float* params, const float* f_in, float* f_out,
                    const float* rho, const float* ux, const float* uy, int ny, int nx, double tau,
                    tlbm_stream_t stream);

/* MSE gradient partials: partials[ceil(n / 64)][nparam + 2] (last columns: loss share, sum |error|);
 * hidden_features * hidden_layers at most 120 */
int tlbm_mlp2d_grad(const tlbm_mlp2d_arch* arch, const float* params, const float* x, const float* t,
                    int64_t n, float* partials, tlbm_stream_t stream);

/* sum the partials in block order, clip the global gradient norm (<= 0: none), one Adam step
 * (lr = 0: only reduce); stats[0] = MSE, stats[1] = sum |error| (stats may be NULL) */
int tlbm_mlp2d_adam(const tlbm_mlp2d_arch* arch, float* params, const float* partials, int nblocks,
                    float* m, float* v, int step, float lr, float beta1, float beta2, float eps,
                    float clip_norm, float* stats, tlbm_stream_t stream);

/* block average by an integer factor (_coarsen_mean, ai/pipeline.py:93-107, in 3-D):
 * out[10][nz/f][ny/f][nx/f] = rho, u_x, u_y, u_z, <uxux>, <uyuy>, <uzuz>, <uxuy>, <uxuz>, <uyuz> */
int tlbm_coarsen_moments(int lattice, int store, const void* f, const int64_t* strides, int nz,
                         int ny, int nx, int factor, float* out, tlbm_stream_t stream);

/* training samples from coarsened moments: periodic central-difference strain
 * (strain_rate_tensor_2d, ai/dataset.py:41-72, in 3-D) as feats[n][7] and
 * nu = max(0, -tau_ij S_ij / (2 S_kl S_kl)), tau_ij = <u_i u_j> - <u_i><u_j>, as labels[n] */
int tlbm_sgs_samples(const float* fields, int nz, int ny, int nx, float* feats, float* labels,
                     tlbm_stream_t stream);

/*
 * Batched lid-driven cavities (benchmark case of many independent cubic D3Q19 BGK cavities).
 * Each cavity of edge L is the reference's cavity step chain on an L^3 grid: pull stream,
 * Zou-He moving lid on the fluid cells of y = L-1 (zou_he_moving_lid_3d, boundaries3d.py:324-404),
 * full-way bounce-back walls x = 0, x = L-1, y = 0, z = 0, z = L-1
 * (bounce_back_cells_3d, boundaries3d.py:112-132; walls of cases/cavity.py:72-79 closed in z),
 * then BGK (collide_bgk3d, solver3d.py:100-104) -- per cavity the same result as tlbm_step
 * with that wall mask and TLBM_BC_LID_Y.  No reference entry point runs a batch in one launch.
 *
 * One buffer holds all cavities: cavity i at element offset `offset`, layout [z][q][y][x]
 * with the (y, x) plane padded to a multiple of 64 elements.  Up to TLBM_CAVITIES_MAX
 * cavities; the table is passed by value to the kernel (no device table).
 */
#define TLBM_CAVITIES_MAX 128

typedef struct tlbm_cavity {
  int32_t edge;         /* L, 3..1024 */
  int32_t reserved;
  double tau;           /* BGK relaxation time (> 0.5) */
  int64_t offset;       /* element offset in the batch buffer (tlbm_cavities_layout) */
  int64_t tiles_before; /* 256-cell tiles of the preceding cavities (tlbm_cavities_layout) */
} tlbm_cavity;

/* host only: fill offset / tiles_before from the edges; totals in elements and tiles */
int tlbm_cavities_layout(tlbm_cavity* cavs, int n, int64_t* total_elements, int64_t* total_tiles);

/* one step of every cavity, f_in -> f_out (distinct buffers of `capacity` elements; store
 * F32, F64, F16_DEV or BF16_DEV).  grid = 0: one workgroup per tile; grid > 0: a persistent
 * grid of that many workgroups striding over the tiles. */
int tlbm_cavities_step(int store, const void* f_in, void* f_out, int64_t capacity, const tlbm_cavity* cavs,
                       int n, double u_lid, int grid, tlbm_stream_t stream);

/* Capture `steps` (even) ping-pong steps a -> b -> a ... into a HIP graph; after a launch the
 * state is back in f_a.  The graph keeps the buffer pointers: free it before the buffers. */
typedef struct tlbm_graph tlbm_graph;
int tlbm_cavities_graph(int store, void* f_a, void* f_b, int64_t capacity, const tlbm_cavity* cavs, int n,
                        double u_lid, int grid, int steps, tlbm_graph** graph);
/* two steps of every cavity in one launch, f_in -> f_out (store F32, F16_DEV or BF16_DEV):
 * the two-step march of tlbm_step2 over 64 x 8 column tiles of every cavity in one grid, with
 * the walls decided from the coordinates -- per cavity the same result as two tlbm_step
 * launches with the wall mask and TLBM_BC_LID_Y, up to rounding of the stored type.
 * zc = 0: the z chunk per workgroup from the launch-shape model; zc > 0: that many planes.
 * Two iterations of the reference's per-cavity step loop (cases/cavity.py) for all cavities. */
int tlbm_cavities_step2(int store, const void* f_in, void* f_out, int64_t capacity, const tlbm_cavity* cavs,
                        int n, double u_lid, int zc, tlbm_stream_t stream);
/* Capture `launches` (even) two-step launches a -> b -> a ... (2 x launches steps) into a HIP
 * graph; after a replay the state is back in f_a. */
int tlbm_cavities_graph2(int store, void* f_a, void* f_b, int64_t capacity, const tlbm_cavity* cavs, int n,
                         double u_lid, int zc, int launches, tlbm_graph** graph);
/* enqueue one replay on `stream`; on the null stream the replay runs on the graph's own stream,
 * fenced by events so that it follows earlier and precedes later null-stream work */
int tlbm_graph_launch(tlbm_graph* graph, tlbm_stream_t stream);
int tlbm_graph_destroy(tlbm_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* TLBM_HIP_H */
