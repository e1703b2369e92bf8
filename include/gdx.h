/* gdx: C-ABI of the MI355X-native GradDFT exchange-correlation grid path.
 *
 * Every pointer argument named d_* or described as "device" is device memory owned by
 * the caller; the library allocates nothing inside these calls.  Work is enqueued on the
 * HIP stream passed as `stream` (a hipStream_t, NULL = default stream) and is
 * asynchronous: results are valid once that stream has reached them.  Every entry point
 * returns GDX_OK (0) or a negative status; gdx_last_error() returns a description of
 * the last failure on the calling thread.
 *
 * Correspondence with the reference (XanaduAI/GradDFT), which has no FFI of its own --
 * each entry point replaces the Python/JAX computation cited beside it.
 */
#ifndef GDX_H_
#define GDX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GDX_OK = 0,
  GDX_ERR_INVALID = -1,     /* bad argument or shape */
  GDX_ERR_HIP = -2,         /* a HIP runtime call failed */
  GDX_ERR_UNSUPPORTED = -3  /* valid request this build does not implement */
};

/* Element type of the matrix products (densities, V_xc, MLP layers).  Per-point functional
 * math, contractions and every reduction over grid points are FP64 in all modes.
 *  GDX_FP64   all products FP64 (v_mfma_f64_16x16x4f64)
 *  GDX_FP32   all products FP32 (v_mfma_f32_16x16x4f32), FP32 AO planes
 *  GDX_BF16X6 as GDX_FP32, but the MLP's hidden h x h layers use split-BF16: every FP32
 *             operand split into three bf16 terms, six products on v_mfma_f32_16x16x32_bf16,
 *             FP32 accumulation (FP32-level accuracy; AO planes FP32). */
enum { GDX_FP64 = 0, GDX_FP32 = 1, GDX_BF16X6 = 2 };

/* Functional kinds evaluated on the grid.
 *  GDX_FUNC_B3LYP        E = sum_r w * (0.8 e_lda + 0.72 e_b88 + 0.19 e_vwn + 0.81 e_lyp)
 *                        (the non-exact-exchange part of grad_dft/popular_functionals.py:375-393)
 *  GDX_FUNC_NEURAL_B3LYP E = sum_r w * sum_f c_theta(x(r))_f e_f(r): the four B3LYP energy
 *                        densities weighted by a coefficient MLP with the head
 *                        c = 2 sigmoid(z / 2) (NeuralFunctional, grad_dft/functional.py:345-419).
 *  GDX_FUNC_B3LYP_NEURAL_CORR  as GDX_FUNC_NEURAL_B3LYP with a linear head around the B3LYP
 *                        coefficients, c_f = c_B3LYP,f + z_f: B3LYP plus a neural correction. */
enum { GDX_FUNC_B3LYP = 0, GDX_FUNC_NEURAL_B3LYP = 1, GDX_FUNC_B3LYP_NEURAL_CORR = 2, GDX_FUNC_LINEAR = 3 };

/* GDX_FUNC_LINEAR: E = sum_r w * clip(sum_f coef[f] e_f(r)) over the four energy-density slots
 * (lda, b88, correlation, lyp) with constant coefficients desc->coef -- the reference's
 * built-in single-part functionals (grad_dft/popular_functionals.py:375-393): LSDA = (1,0,0,0),
 * B88 = (1,1,0,0) (b88_density is [lda, b88] under a unit coefficient), VWN = (0,0,1,0) with
 * GDX_CORR_VWN5, PW92 = (0,0,1,0) with GDX_CORR_PW92, LYP = (0,0,0,1).
 * desc->correlation selects the correlation slot for every kind: VWN5 (vwn_c_e, :141-195) or
 * PW92 (pw92_c_e, :105-139). */
enum { GDX_CORR_VWN5 = 0, GDX_CORR_PW92 = 1 };

/* Shape and functional description of one XC evaluation. */
typedef struct gdx_xc_desc {
  int32_t npts;            /* grid points N */
  int32_t npts_pad;        /* rows of the AO / weight buffers, multiple of 64, >= npts */
  int32_t nao;             /* basis functions n */
  int32_t nao_pad;         /* columns of the AO buffer, multiple of 16, >= nao */
  int32_t spin_restricted; /* 1 when rdm1[0] == rdm1[1]: densities are formed once */
  int32_t kind;            /* GDX_FUNC_* */
  int32_t precision;       /* GDX_FP64 | GDX_FP32 | GDX_BF16X6 */
  int32_t mlp_nc;          /* MLP inputs (6: rho, |grad rho|^2, tau per spin) */
  int32_t mlp_h;           /* hidden width */
  int32_t mlp_nhidden;     /* hidden layers (dense + ELU each) */
  int32_t mlp_nd;          /* outputs = energy densities (4) */
  int32_t correlation;     /* GDX_CORR_VWN5 | GDX_CORR_PW92 */
  double coef[4];          /* GDX_FUNC_LINEAR coefficients (ignored by the other kinds) */
} gdx_xc_desc;

const char* gdx_last_error(void);
int gdx_version(void);

/* AO values on the grid: replaces PySCF numint.eval_ao(deriv=1) + ao_grads(order=2) as
 * packaged by grad_dft/interface/pyscf.py:824-827, 677-705.
 * Inputs (device): shell_i32 (nshell,6) {l, nprim, prim_offset, ao_offset, nfunc,
 * tmat_offset}, shell_f64 (nshell,3) centres, prim_f64 (nprim,2) {exponent, coef},
 * tmat_f64 component matrices, coords (npts,3).
 * Output (device): ao, 5 planes of (npts_pad, nao_pad) doubles: phi, dphi/dx, dphi/dy,
 * dphi/dz, laplacian(phi) (= the sum of grad_n_ao[2]'s xx, yy, zz).  Padding is zeroed. */
int gdx_ao_eval(void* stream, int32_t nshell, const int32_t* shell_i32, const double* shell_f64,
                const double* prim_f64, const double* tmat_f64, int32_t npts, int32_t npts_pad,
                const double* coords, int32_t nao, int32_t nao_pad, double* ao);

/* Bytes of device workspace gdx_xc_forward / gdx_xc_backward need for `desc`. */
int gdx_xc_workspace_size(const gdx_xc_desc* desc, size_t* bytes);

/* E_xc = Functional.xc_energy(params, grid, cinputs, densities) of
 * grad_dft/functional.py:219-253 with densities and coefficient inputs formed from rdm1
 * as in compute_densities / compute_coefficient_inputs (:160-217) and
 * grad_dft/molecule.py:388-502.
 * ao: gdx_ao_eval output in the element type of desc->precision (double planes for
 * GDX_FP64, float planes otherwise); weights (npts_pad) grid weights, zero past npts;
 * rdm1 (2, nao, nao); params: flat MLP parameters [W0 (nc,h), b0 (h), W1 (h,h), b1, ...,
 * W_last (h,nd), b_last] (NULL for GDX_FUNC_B3LYP).  exc: one device double.
 * The workspace keeps what gdx_xc_backward needs; do not touch it in between. */
int gdx_xc_forward(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
                   const double* rdm1, const double* params, double* exc, void* workspace);

/* Autotuner (no counterpart in the reference): times the candidate output tiles (64, 128) of
 * the V_xc build for this descriptor's shape on the device, with the given AO planes and
 * workspace (whose contents it overwrites), and records the faster one for every later call
 * with the same shape (N_pad, n_pad, spins, precision) in this process; a shape measured
 * before keeps its choice.  vt_inout (may be NULL): 0 on entry measures, 64 or 128 pins that
 * tile without measuring; on exit the first-order choice.  Synchronises the stream; not for
 * use under graph capture. */
int gdx_xc_autotune(void* stream, const gdx_xc_desc* desc, const void* ao, void* workspace, int32_t* vt_inout);

/* Custom coefficient functions (the reference's Functional(coefficients=..., ...) with a user
 * module, grad_dft/functional.py:160-253): the library forms the four B3LYP energy densities
 * e_f (lda, b88, correlation, lyp; unclipped) and the six coefficient inputs x_j =
 * log(|feature| + 1e-4) of (rho_up, rho_dn, |grad rho|^2 up/dn, tau up/dn) per point; the caller
 * evaluates its coefficients and E = sum_r w clip(sum_f c_f clip(e_f)) itself, and hands the
 * cotangents back for dE/dD.  desc: kind GDX_FUNC_NEURAL_B3LYP (its MLP fields only size the
 * workspace), mlp_nc 6, GDX_FP64.
 * gdx_xc_custom_inputs: e_out (npts, 4), x_out (npts, 6) (device), from rdm1 (2, nao, nao).
 * gdx_xc_custom_vxc: d_rdm1 (2, nao, nao) = sum_r (g_e . de/dD + g_x . dx/dD) for cotangents
 *   g_e (npts, 4), g_x (npts, 6); reads the densities the preceding gdx_xc_custom_inputs call
 *   left in the same workspace. */
int gdx_xc_custom_inputs(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
                         const double* rdm1, void* workspace, double* e_out, double* x_out);
int gdx_xc_custom_vxc(void* stream, const gdx_xc_desc* desc, const void* ao, const double* g_e, const double* g_x,
                      void* workspace, double* d_rdm1);

/* Grid densities per spin, (npts_pad, 2, 6) doubles: rho, d rho/dx, d rho/dy, d rho/dz,
 * tau, laplacian -- Molecule.density / grad_density / kinetic_density / lapl_density of
 * grad_dft/molecule.py:388-502 in one pass.  Uses (and overwrites) the workspace. */
int gdx_xc_densities(void* stream, const gdx_xc_desc* desc, const void* ao, const double* rdm1,
                     double* dens, void* workspace);

/* Vector-Jacobian product of gdx_xc_forward for the cotangent g_exc (one device double):
 * d_rdm1 (2, nao, nao) = g * dE_xc/dD_s -- the unsymmetrised Fock XC term of
 * grad_dft/train.py:86-121 (value_and_grad w.r.t. rdm1) -- and d_params (same layout as
 * params) = g * dE_xc/dtheta.  Either output may be NULL to skip it. */
int gdx_xc_backward(void* stream, const gdx_xc_desc* desc, const void* ao,
                    const double* weights, const double* rdm1, const double* params,
                    const double* g_exc, double* d_rdm1, double* d_params, void* workspace);

/* Second-order pass for reverse mode through an SCF loop.  With E(theta, D) = E_xc and a
 * direction dir (2, nao, nao), forms Phi = g E + <dir, dE/dD> (g = *g_exc) and returns
 * exc = E, edot = <dir, dE/dD> (forward-mode derivative of E along dir), d_rdm1 = dPhi/dD
 * (2, nao, nao) and d_params = dPhi/dtheta.  In the reference this is what jax.grad
 * computes when it differentiates the Fock matrix's XC term (train.py:86-121) inside
 * diff_scf_loop (grad_dft/evaluate.py:917-1038): the cotangent of that term is `dir`, the
 * cotangent of E_xc is g.  exc, edot, d_rdm1, d_params may each be NULL.  workspace:
 * gdx_xc_hvp_workspace_size bytes (its own buffer; independent of gdx_xc_forward's). */
int gdx_xc_hvp_workspace_size(const gdx_xc_desc* desc, size_t* bytes);
int gdx_xc_hvp(void* stream, const gdx_xc_desc* desc, const void* ao, const double* weights,
               const double* rdm1, const double* params, const double* dir, const double* g_exc,
               double* exc, double* edot, double* d_rdm1, double* d_params, void* workspace);

/* ---- A batch of molecules ----------------------------------------------------------
 * The XC evaluation of several molecules (one functional and precision, different grids and
 * bases) as one: per-point buffers concatenate the molecules' rows, the MLP GEMMs run once
 * over all rows, and the AO-side kernels (density, V_xc, their padding and reduction) are
 * single launches over a ragged queue of (molecule, tile) work items.  This replaces the
 * molecule-by-molecule loop of the reference's losses (grad_dft/train.py:519-527) for the XC
 * part of every SCF cycle.  Densities, their cotangents and outputs are concatenated: molecule
 * m's (2, n_m, n_m) block starts at dm_off[m] (gdx_xc_batch_offsets, nmol + 1 entries). */
typedef struct gdx_xc_batch_mol {
  const void* ao;          /* gdx_ao_eval planes (element type of the precision), npts_pad x nao_pad */
  const double* weights;   /* (npts_pad) grid weights, zero past npts */
  int32_t npts, npts_pad;  /* npts_pad a multiple of 256 */
  int32_t nao, nao_pad;
} gdx_xc_batch_mol;
/* desc: kind, precision, spin_restricted (all molecules), MLP and correlation fields; its size
 * fields are ignored.  The plan allocates its workspaces here (nothing is allocated by the
 * calls below) and copies the weights; the AO planes are referenced and must outlive it. */
int gdx_xc_batch_create(const gdx_xc_desc* desc, int32_t nmol, const gdx_xc_batch_mol* mols, int32_t with_hvp,
                        void** plan);
int gdx_xc_batch_destroy(void* plan);
int gdx_xc_batch_offsets(const void* plan, int64_t* dm_off);
/* exc (nmol) = E_xc of every molecule (gdx_xc_forward per molecule). */
int gdx_xc_batch_forward(void* stream, void* plan, const double* rdm1, const double* params, double* exc);
/* With cotangents g (nmol): d_rdm1 = g_m dE_m/dD_m (concatenated), d_params = sum_m g_m dE_m/dtheta.
 * Uses the state of the last gdx_xc_batch_forward on the plan. */
int gdx_xc_batch_backward(void* stream, void* plan, const double* params, const double* g_exc, double* d_rdm1,
                          double* d_params);
/* Second-order pass of every molecule (gdx_xc_hvp): Phi = sum_m g_m E_m + <dir_m, dE_m/dD_m>;
 * exc, edot (nmol each), d_rdm1 (concatenated), d_params; plan created with with_hvp. */
int gdx_xc_batch_hvp(void* stream, void* plan, const double* rdm1, const double* params, const double* dir,
                     const double* g_exc, double* exc, double* edot, double* d_rdm1, double* d_params);

/* ---- SCF side ---------------------------------------------------------------------
 * eri: 8-fold packed (ij|kl), pair index I = i(i+1)/2 + j (i >= j), element I(I+1)/2 + K
 * (I >= K); every matrix row-major, rdm1 (2, n, n). */

/* Coulomb matrix J (n, n) of D_up + D_dn: coulomb_potential(rdm1.sum(0), rep_tensor) of
 * grad_dft/molecule.py:108-115, 788-811, from the packed ERI (each element read once).
 * workspace: gdx_coulomb_workspace_size bytes. */
int gdx_coulomb_workspace_size(int32_t n, size_t* bytes);
int gdx_coulomb_packed(void* stream, int32_t n, const double* eri, const double* rdm1, double* J,
                       void* workspace);

/* gdx_coulomb_packed without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The kernels use a fixed 32 MiB device array
 * of this library (one per GPU, held until the process ends) in place of the workspace, so
 * the call allocates nothing; n above gdx_coulomb_noalloc_max_n (214) is rejected with
 * GDX_ERR_INVALID.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (different streams, or host threads without ordering) race and must not be issued. */
int gdx_coulomb_noalloc_max_n(int32_t* n_max);
int gdx_coulomb_packed_noalloc(void* stream, int32_t n, const double* eri, const double* rdm1, double* J);

/* Non-XC energy e_nuc + sum D h1e + 1/2 sum D J (D = D_up + D_dn) into one device double:
 * nonXC of grad_dft/molecule.py:694-783.  With K0 (2, n, n) non-NULL the exact-exchange energy
 * of B3LYP from the initial density, -(a0/2) sum_s sum D_s K0_s, is added (its grid form is
 * HF_energy_density, molecule.py:507-541, integrated in xc_energy). */
int gdx_nonxc_energy(void* stream, int32_t n, const double* rdm1, const double* h1e, const double* J,
                     const double* K0, double a0, double e_nuc, double* out);

/* Fock matrix of predict (grad_dft/train.py:148-215): per spin F = clip(h1e + J + vxc_s),
 * symmetrise, clip, add the explicit exchange term X_s + X_s^T with X_s = -(a0/2) K0_s when
 * K0 is non-NULL, clip; clip(x) = x if |x| > 1e-30 else 0.  vxc: gdx_xc_backward's d_rdm1. */
int gdx_fock_assemble(void* stream, int32_t n, const double* h1e, const double* J, const double* vxc,
                      const double* K0, double a0, double* fock);

/* Exchange K_s[a][c] = sum_bd (ab|cd) D_s[b][d] for nspin matrices (nspin, n, n), D general
 * (not necessarily symmetric): the exact-exchange term of B3LYP from the initial density
 * (stands in for the semi-numerical chi term, grad_dft/molecule.py:507-613,
 * interface/pyscf.py:1067-1126), or from the current one.  Each element is summed in a fixed
 * order (reproducible bit for bit); K must not alias D. */
int gdx_exchange_packed(void* stream, int32_t n, int32_t nspin, const double* eri, const double* D, double* K);

/* The same K_s from an exchange-ordered copy of the ERI, read once per spin at streaming rate:
 * over ordered pairs I = a n + c the matrix M[I][b n + d] = (ab|cd) is symmetric, and
 * K_s = M vec(D_s) is the packed symmetric product of gdx_coulomb_packed with P' = n^2.
 * gdx_exchange_xo_size: bytes of the copy, n^2 (n^2 + 1) / 2 doubles (4x the packed ERI;
 * water def2-TZVP 14 MB, naphthalene def2-SVP 4.2 GB).  gdx_exchange_xo_build: the copy from the
 * 8-fold packed ERI (once per molecule).  gdx_exchange_xo: K (nspin, n, n) from it, D general;
 * workspace gdx_exchange_xo_workspace_size bytes.  Summation order fixed (reproducible bit for
 * bit); agrees with gdx_exchange_packed to rounding.  Replaces the same reference term as
 * gdx_exchange_packed (grad_dft/molecule.py:507-613). */
int gdx_exchange_xo_size(int32_t n, size_t* bytes);
int gdx_exchange_xo_build(void* stream, int32_t n, const double* eri, double* xo);
int gdx_exchange_xo_workspace_size(int32_t n, size_t* bytes);
int gdx_exchange_xo(void* stream, int32_t n, int32_t nspin, const double* xo, const double* D, double* K,
                    void* workspace);

/* S = L L^T and Linv = L^-1 (the cholesky / inv of safe_general_eigh,
 * grad_dft/utils/eigenproblem.py:125-126).  status: device int, 1 if S is not positive definite. */
int gdx_cholesky_inv(void* stream, int32_t n, const double* S, double* L, double* Linv, int32_t* status);

/* Batched symmetric eigensolver (jnp.linalg.eigh in safe_eigh, eigenproblem.py:26-38):
 * A (nbatch, n, n) -> w (nbatch, n) ascending, V (nbatch, n, n) eigenvectors as columns.
 * V0 (nbatch, n, n), optional (NULL = none): orthonormal starting eigenvector estimates (the
 * previous SCF cycle's), which cut the Jacobi sweeps for a nearby matrix; it must not alias
 * A or V.  workspace: gdx_eigh_workspace_size bytes (used when 3 n^2 doubles exceed the
 * device's LDS per workgroup less 13 KB: n > 79 on gfx950's 160 KB).  sweeps: optional
 * device int per matrix (Jacobi sweeps used). */
int gdx_eigh_workspace_size(int32_t nbatch, int32_t n, size_t* bytes);
int gdx_eigh_batched(void* stream, int32_t nbatch, int32_t n, const double* A, const double* V0,
                     double* w, double* V,
                     void* workspace, int32_t max_sweeps, double tol, int32_t* sweeps);

/* One SCF diagonalisation step per spin, forward only (the body of diff_scf_loop's cycle
 * after the DIIS coefficients, grad_dft/evaluate.py:989-1005 with cdiis_minimize's
 * extrapolation :1198-1205, safe_fock_solver / safe_general_eigh utils/eigenproblem.py:110-149
 * and make_rdm1 molecule.py:816-889): F = sum_i x[s][1+i] focks[i][s] (focks (m, nspin, n, n),
 * x (nspin, m+1) from gdx_diis_coefficients), C' = sym(Linv F Linv^T) = V diag(w) V^T
 * (Jacobi as gdx_eigh_batched, warm-started from V0 when given), C = Linv^T V and
 * rdm1 = C diag(occ) C^T with occ (nspin, n).  Outputs F, w, V (orthonormal eigenvectors, the
 * next call's V0), C, rdm1.  workspace: gdx_scf_step_workspace_size bytes (used when 4 n^2
 * doubles exceed the device's LDS per workgroup less 13 KB: n > 68 on gfx950's 160 KB). */
int gdx_scf_step_workspace_size(int32_t nspin, int32_t n, size_t* bytes);
int gdx_scf_step(void* stream, int32_t nspin, int32_t n, int32_t m, const double* focks, const double* x,
                 const double* Linv, const double* V0, const double* occ, double* F, double* w, double* V,
                 double* C, double* rdm1, void* workspace, int32_t max_sweeps, double tol, int32_t* sweeps);

/* Status word of an SCF cycle (no counterpart in the reference, whose checks are host
 * asserts: charge conservation and NaN-free DIIS, grad_dft/evaluate.py:212-217, 506-508, 1384).
 * OR-ed into the device int *status, so a whole captured loop reports once:
 *  GDX_STATUS_EIGH_NOT_CONVERGED  a Jacobi eigensolve reached max_sweeps (sweeps: device int per
 *                                 spin from gdx_scf_step / gdx_eigh_batched, may be NULL)
 *  GDX_STATUS_NONFINITE_ENERGY    *energy (device double, may be NULL) is NaN or infinite
 *  GDX_STATUS_CHARGE              |tr(D_s S) - sum occ_s| > tol for a spin (rdm1 (nspin, n, n),
 *                                 S (n, n), occ (nspin, n)) */
enum { GDX_STATUS_EIGH_NOT_CONVERGED = 1, GDX_STATUS_NONFINITE_ENERGY = 2, GDX_STATUS_CHARGE = 4 };
int gdx_scf_status(void* stream, int32_t nspin, int32_t n, const double* rdm1, const double* S, const double* occ,
                   const int32_t* sweeps, int32_t max_sweeps, const double* energy, double tol, int32_t* status);

/* x_s = B_s^-1 e_0 for nspin bordered DIIS matrices B (nspin, m1, m1), m1 <= 16:
 * JittableDiis.cdiis_minimize, grad_dft/evaluate.py:1198-1205. */
int gdx_diis_solve(void* stream, int32_t nspin, int32_t m1, const double* B, double* x);

/* DIIS error of a cycle, JittableDiis.update (grad_dft/evaluate.py:1111-1123):
 * err_s = F_s D_s S - (F_s D_s S)^T for nspin n x n matrices (S shared).  workspace:
 * gdx_diis_error_workspace_size bytes. */
int gdx_diis_error_workspace_size(int32_t nspin, int32_t n, size_t* bytes);
int gdx_diis_error(void* stream, int32_t nspin, int32_t n, const double* F, const double* D, const double* S,
                   double* err, void* workspace);

/* JittableDiis coefficients (grad_dft/evaluate.py:1185-1205) in one call: errs (m, nspin,
 * n2) holds the m stored error vectors FDS - SDF (n2 = n^2 elements each); the Gram matrix per
 * spin, the bordered matrix B (nspin, m+1, m+1) -- +1 borders for the slots i <= cycle, the
 * Gram diagonal there and 1 on the diagonal of the others, B[0][0] = 0 -- and x = B^-1 e_0
 * (nspin, m+1) are written.  m + 1 <= 16. */
int gdx_diis_coefficients(void* stream, int32_t nspin, int32_t m, int64_t n2, const double* errs, int32_t cycle,
                          double* B, double* x);

/* ---- Electron-repulsion integrals on the device ---------------------------------------
 * The packed ERI of a basis (HOST shell arrays in the gdx_ao_eval packing, shells up to d):
 * mol.intor("int2e") as called by calc_eri_with_pyscf, grad_dft/interface/pyscf.py:746-803,
 * in the 8-fold packed layout of the SCF entry points (P(P+1)/2 doubles, P = n(n+1)/2).
 * The plan (primitive-pair expansions and the shell-quartet lists, grouped by angular momenta)
 * is built and uploaded once; gdx_eri_packed writes every element (no clearing needed) and is
 * reproducible.  Primitive pairs with (2 pi / p)|c_a c_b| exp(-ab/p |A-B|^2) < screen are
 * dropped (0 keeps all).  GDX_ERR_UNSUPPORTED for a basis with f shells. */
int gdx_eri_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                        const double* tmat_f64, int32_t nao, double screen, void** plan);
int gdx_eri_plan_destroy(void* plan);
int gdx_eri_plan_info(const void* plan, int64_t* nquartets, int64_t* nentries);
int gdx_eri_packed(void* stream, const void* plan, double* eri);

/* ---- One-electron integrals and grid weights on the device -----------------------------
 * S (overlap), T (kinetic) and V (nuclear attraction, -sum_C Z_C <a|1/|r-C||b>): mf.get_ovlp()
 * and mf.get_hcore() = T + V as read by _package_outputs, grad_dft/interface/pyscf.py:828-829.
 * The plan takes HOST arrays in the gdx_ao_eval packing (shells up to f; ntmat = doubles in
 * tmat_f64), the nuclear charges (natm) and positions (natm, 3; bohr), and uploads the
 * primitive-pair data once.  gdx_int1e_eval writes every element of the device (nao, nao)
 * outputs S, T, V (no clearing needed); results are reproducible. */
int gdx_int1e_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                          const double* tmat_f64, int32_t ntmat, int32_t nao, int32_t natm, const double* charges,
                          const double* atom_xyz, void** plan);
int gdx_int1e_plan_destroy(void* plan);
int gdx_int1e_eval(void* stream, const void* plan, double* S, double* T, double* V);

/* Becke fuzzy-cell weights of an atom-centred grid: the partition behind the weights of PySCF's
 * dft.Grids that grid_from_pyscf takes (grad_dft/interface/pyscf.py:54-60).  Device inputs:
 * coords (npts, 3), owner (npts; the atom whose radial x angular rule made the point), vol
 * (npts; that rule's quadrature volume), atoms (natm, 3), adist (natm, natm; |R_i - R_j|),
 * adjust (natm, natm; the antisymmetric size adjustment a_ij).  Output (device): weights[r] =
 * vol[r] P_owner(r) / sum_i P_i(r) with P_i = prod_{j != i} s(nu_ij), nu_ij = mu_ij +
 * a_ij (1 - mu_ij^2), mu_ij = (|r - R_i| - |r - R_j|) / |R_i - R_j|, s = (1 - f(f(f(nu)))) / 2,
 * f(x) = (3 - x^2) x / 2, evaluated in a fixed order without contraction (equal to the same
 * formulas evaluated sequentially in double precision).  adist/adjust may be NULL when natm = 1. */
int gdx_becke_weights(void* stream, int32_t npts, const double* coords, const int32_t* owner, const double* vol,
                      int32_t natm, const double* atoms, const double* adist, const double* adjust, double* weights);

/* ---- Semi-numerical exact exchange ----------------------------------------------------
 * nu_ab(r) = int dr' phi_a(r') f(|r - r'|) phi_b(r') at grid points r (f = 1/|r-r'|, or
 * erf(omega |r-r'|)/|r-r'| for omega > 0): PySCF's int1e_grids_sph as called by
 * grad_dft/external/_hf_density.py:18-67.  Device layouts: u, chi, v are (nspin, nao, npts_pad)
 * (orbital-major, grid points contiguous); the reference's chi (N, omega, 2, n) is a permuted
 * view of one (2, n, npts_pad) array per omega. */

/* Plan for a basis (HOST arrays in the gdx_ao_eval packing, shells up to f): the Hermite
 * expansion of every primitive pair of every ordered shell pair, folded with the contraction
 * coefficients and component matrices, uploaded once.  Primitive pairs whose prefactor
 * (2 pi / p) |c_a c_b| exp(-ab/p |A-B|^2) is below `screen` are dropped (0 keeps all).  The
 * plan allocates device memory here, never in gdx_hf_chi; free it with gdx_hf_plan_destroy. */
int gdx_hf_plan_create(int32_t nshell, const int32_t* shell_i32, const double* shell_f64, const double* prim_f64,
                       const double* tmat_f64, int32_t nao, double screen, void** plan);
int gdx_hf_plan_destroy(void* plan);
int gdx_hf_plan_info(const void* plan, int64_t* nentries, int64_t* mdoubles);

/* chi[s][b][r] = sum_a u[s][a][r] nu_ab(r) for r < npts (zero beyond): with u = D0_s^T phi(r)
 * (gdx_ao_matmul of the initial density) this is generate_chi_tensor's
 * chi[r, omega, s, b] = sum_ad D0_s[a][d] phi_a(r) nu_db(r), grad_dft/interface/pyscf.py:1067-1126.
 * coords (npts, 3) bohr.  Bitwise reproducible (no atomics). */
int gdx_hf_chi(void* stream, const void* plan, int32_t npts, int32_t npts_pad, const double* coords, int32_t nspin,
               const double* u, double omega, double* chi);

/* out[s][c][r] = sum_a phi[r][a] M_s[a][c]: phi is the AO value plane (npts_pad, nao_pad) of
 * gdx_ao_eval, M (nspin, nao, nao).  With M = D_s: the orbital-side factor of chi (D0) and of
 * HF_energy_density (current D), grad_dft/molecule.py:507-541. */
int gdx_ao_matmul(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nao_pad, const double* phi,
                  int32_t nspin, const double* M, double* out);

/* e[s][r] = -1/2 sum_c chi[s][c][r] v[s][c][r] with v = gdx_ao_matmul(D): HF_energy_density,
 * grad_dft/molecule.py:507-541 (one omega).  e (nspin, npts_pad). */
int gdx_hf_energy_density(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nspin, const double* chi,
                          const double* v, double* e);

/* out[s][a][b] = scale * sum_r g_s(r) phi_a(r) chi[s][b][r], g_s = g + s * g_spin_stride: with
 * g = gr (dE_xc / d e_HF per point) and scale = -1/2 this is HF_density_grad_2_Fock,
 * grad_dft/molecule.py:545-613 (one omega); with g = the grid weights and scale = 1 it is the
 * semi-numerical exchange matrix K~ of B3LYP's constant coefficient.  Partial sums over point
 * chunks are reduced in a fixed order (bitwise reproducible).  workspace:
 * gdx_hf_fock_workspace_size bytes. */
int gdx_hf_fock_workspace_size(int32_t npts, int32_t nao, int32_t nspin, size_t* bytes);
int gdx_hf_fock(void* stream, int32_t npts, int32_t npts_pad, int32_t nao, int32_t nao_pad, const double* phi,
                int32_t nspin, const double* chi, const double* g, int64_t g_spin_stride, double scale, double* out,
                void* workspace);

/* Kernel timing (no counterpart in the reference): while on, every kernel this library
 * enqueues is bracketed by HIP events on its own stream.  gdx_profile_end waits for them and
 * returns per-kernel-name total milliseconds and launch counts (names: 64 chars each). */
int gdx_profile_begin(void);
int gdx_profile_end(int32_t max_entries, char* names, double* total_ms, int32_t* counts,
                    int32_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* GDX_H_ */
