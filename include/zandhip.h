/* zandhip: MI355X (gfx950) propagation engine for the auxiliary-mode-expansion
 * (AME) equations of motion and their DOP54 adaptive Runge-Kutta driver.
 *
 * C ABI.  Complex arrays are interleaved (re, im) doubles in the reference's
 * array layout (numpy complex128, C order):
 *   sigma (NK,n,n)   psi (NK,NL,NX,NOT,n)   omega (NK,NL,NX,NOT,NL,NX,NOT)
 *   mode constants (NK,NL,NX,NOT)
 * Host pointers are borrowed for the duration of a call only; device memory is
 * owned by the problem handle.  Every call returns 0 on success and a negative
 * ZH_E* code on failure; zh_last_error() gives the message (thread-local).
 * A handle is not thread-safe: the caller serialises calls on one handle;
 * distinct handles may be used concurrently.
 */
#ifndef ZANDHIP_H
#define ZANDHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ZH_ABI_VERSION 1
#define ZH_MAX_LEADS 8
#define ZH_MAX_DRIVE 8

enum { ZH_FP64 = 0, ZH_MIXED = 1 };
enum { ZH_OK = 0, ZH_EINVAL = -1, ZH_EHIP = -2, ZH_ENOMEM = -3, ZH_ESTATE = -4, ZH_ENONFINITE = -5,
       ZH_ELIMIT = -6 };
/* bias waveform kinds (Zandpack/Pulses.py:52-94 shapes) */
enum { ZH_PULSE_NONE = 0, ZH_PULSE_STEP = 1, ZH_PULSE_GAUSS = 2, ZH_PULSE_RAMP = 3, ZH_PULSE_SIN = 4,
       ZH_PULSE_BOX = 5, ZH_PULSE_GENERIC = 6, ZH_PULSE_STAIRS = 7, ZH_PULSE_SPLINE = 8 };
/* time factors of the drive terms of zh_set_drive */
enum { ZH_WAVE_CONST = 0, ZH_WAVE_COS_WINDOW = 1, ZH_WAVE_SIN_FERMI = 2, ZH_WAVE_GAUSS_SIN = 3 };
/* dissipator rate kinds */
enum { ZH_DISS_NONE = 0, ZH_DISS_CONST = 1, ZH_DISS_STEP = 2 };
/* time bookkeeping after an accepted step */
enum { ZH_TIME_REFERENCE = 0, /* t += h_next, as Zandpack/mpi/zand:406-407,427 */
       ZH_TIME_CORRECT = 1 }; /* t += h of the step just taken */

typedef struct zh_problem zh_problem;

typedef struct {
  int NK, n, NL, NX, NOT; /* k points (or spin), orbitals, leads, pole slots, channels */
  int precision;          /* ZH_FP64 or ZH_MIXED */
  double hbar;            /* eV*fs; Zandpack/td_constants.py:3 uses 0.6582119569 */
  int device;             /* HIP device ordinal */
} zh_dims;

typedef struct {
  double t0, h, eps; /* in: step start, step size, absolute error tolerance */
  int force_accept;  /* in: 1 = accept regardless of TE (fixed-step / replay) */
  double te;         /* out: sqrt(sum |sum_j CT_j k_j|^2) over sigma, Psi, Omega */
  double h_next;     /* out: 0.9 h (eps/TE)^(1/5) */
  int accepted;      /* out */
} zh_attempt_info;

/* Version / diagnostics. */
int zh_abi_version(void);
const char* zh_last_error(void);

/* Replaces the problem set-up done by Zandpack/mpi/zand:113-132, 568-678
 * (array shapes, NOT cut, orbital pruning, constant products). */
int zh_problem_create(const zh_dims* dims, zh_problem** out);
void zh_problem_destroy(zh_problem* p);

/* A problem of `instances` independent propagations of the same shape (dims->NK k
 * blocks each) stacked on the k axis, so that every kernel of an attempt runs once for
 * all of them (one launch per kernel class across the instances).  Every other call
 * then speaks NK = instances * dims->NK: constants, state and currents of instance g are
 * its k blocks [g dims->NK, (g+1) dims->NK) (zh_instances reports the count).  Each
 * instance has its own time, step, tolerance, pulse (zh_instance_pulse), error norm and
 * acceptance: zh_attempt_instances.  Results of an instance are bitwise those of the same
 * instance alone on the same tiles.  The instances share the device mode order (the
 * union of their couplings), the Hubbard U, drive and dissipator.  No counterpart in the
 * reference (one problem per mpirun, Zandpack/mpi/zand:297-431); for the batched
 * workload of uneven instances, grouped by shape. */
int zh_problem_create_instances(const zh_dims* dims, int instances, zh_problem** out);
int zh_instances(zh_problem* p, int* instances);

/* Device mode order, chosen before zh_set_constants: 0 auto (default), 1 the
 * reference order, 2 active modes (dGP or dGM non-zero) first, padded to a
 * 64-mode tile, then the passive Fermi modes.  Order 2 lets the Omega kernels skip
 * the vanishing source terms per tile and the passive x passive block of Omega,
 * which stays exactly zero when it starts at zero (the reference's Fermi-block
 * pruning, Zandpack/mpi/nozand:818-890).  Auto picks 2 when the padding costs at
 * most 10% of the modes.  Results do not depend on the order. */
int zh_set_layout(zh_problem* p, int mode);
/* out[5] = {active-first order in use, passive block skipped, device modes,
 * active block size, coupled orbitals}. */
int zh_layout_info(zh_problem* p, int* out);

/* How the two H products of the right-hand side run (H sigma,
 * Zandpack/mpi/zand:340-343, and psi H^T, Zandpack/mpi/zand:746-778, both dense
 * matrix products there): 0 auto (default), 1 dense MFMA products, 2 gathers over
 * a row-compressed copy of H (at most 64 non-zeros per row; ZH_EINVAL otherwise).
 * Auto takes 2 when every row holds at most 64 non-zeros and at most n/8.  The
 * choice is made again whenever H is installed.  Results agree to rounding. */
int zh_set_sparse(zh_problem* p, int mode);
/* out[3] = {H real, row-compressed H products in use, entries per compressed row}. */
int zh_hamiltonian_info(zh_problem* p, int* out);

/* Constants, reference layout, complex128 interleaved unless noted:
 *   H0 (NK,n,n); xi, Ixi (NK,NL,NX,NOT,n); GL_P, dGP=GG_P-GL_P, dGM=GG_M-GL_M,
 *   Xpp, Xpm (NK,NL,NX,NOT).
 * Replaces the Arrays/ loads of Zandpack/mpi/zand:218, 568-576. */
int zh_set_constants(zh_problem* p, const double* H0, const double* xi, const double* Ixi,
                     const double* GL_P, const double* dGP, const double* dGM, const double* Xpp,
                     const double* Xpm);

/* Device-evaluated H(t, sigma) = H0 + U diag(sigma[partner]_ii - n0[partner]_i) for NK = 2
 * spin blocks (examples/Hubbard/Calculation/Bias.py:13-20).  U = 0 disables it.
 * n0 is (NK, n) real.  With an overlap (zh_set_overlap) the term enters as nozand applies
 * dH (S^-1 (H0 + dH), Zandpack/mpi/nozand:427, 456-461, 1075-1085). */
int zh_set_hubbard(zh_problem* p, double U, const double* n0);

/* Drive terms of the device-evaluated Hamiltonian: H(t, sigma) = H0 + (Hubbard term) +
 * sum_i f_i(t) M_i, the time-dependent dH(t, sigma) a run's Bias.py hands the propagator
 * every stage (Zandpack/mpi/zand:318-319; mpi/nozand:427), e.g. a Rabi drive
 * (examples/nb_tutorial_updated/Tight-binding/T1_Rabi/Bias.py:16-24), switched site
 * potentials (examples/Reproduce_and_checks/scipy_vs_ARK4_Zand/shared.py:4-19) or a
 * dipole ramp (examples/nb_tutorial_updated/Tight-binding/T2_TB_Chain/Bias.py:21-25).
 * nterms <= ZH_MAX_DRIVE terms; term i has the time factor kinds[i] (ZH_WAVE_*) with
 * amps[i] and par[4 i .. 4 i + 3]:
 *   CONST       amp
 *   COS_WINDOW  amp cos(w t + phi) for t_on < t <= t_off, else 0;   par = {w, phi, t_on, t_off}
 *   SIN_FERMI   amp sin(w t + phi) / (1 + exp((t - t_off)/width));  par = {w, phi, t_off, width}
 *   GAUSS_SIN   amp exp(-((t-d)/s)^2) sin(2 pi f (t-d) - phi);       par = {s, f, phi, d}
 * and the nnz[i] entries (ks, rows, cols, vals) of M_i (vals complex128 interleaved),
 * listed term after term.  nterms = 0 removes the drive.  A row-compressed H then holds
 * the union of H0's and the drive's entries. */
int zh_set_drive(zh_problem* p, int nterms, const int* kinds, const double* amps,
                 const double* par, const int* nnz, const int* ks, const int* rows,
                 const int* cols, const double* vals);

/* Device-evaluated bias Delta_a(t) = weights[a] * V(t) (Bias.bias(t, a) hook,
 * Zandpack/mpi/zand:320-321).  kind: ZH_PULSE_*; amp, p1, p2 per kind:
 *   STEP amp for t >= p1; GAUSS amp exp(-((t-p1)/p2)^2); RAMP amp (1 - exp(-t/p1));
 *   SIN amp sin(p1 t / hbar); BOX amp (tanh(t/p2) - tanh((t-p1)/p2))/2. */
int zh_set_pulse(zh_problem* p, int kind, double amp, double p1, double p2, const double* weights);

/* Any pulse kind with up to four parameters par[0..3] (zh_set_pulse is the two-
 * parameter form):
 *   GENERIC amp exp(-((t-d)/s)^2) sin(2 pi f (t-d) - phase), par = {s, f, phase, d}
 *           (generic_pulse with the Gaussian envelope env1, Zandpack/Pulses.py:52-66)
 *   STAIRS  amp sum_{i=1}^{n-1} (1 - 1/(1+exp((t - i w)/soft))), par = {w, n, soft}
 *           (stairs, Pulses.py:83-88)
 *   SPLINE  amp S(t) (tanh((t-a)/ts) - tanh((t-b)/ts))/2 inside the knots, 0 outside,
 *           par = {a, b, ts}; S the piecewise cubic of zh_set_pulse_spline (the
 *           data-driven air_photonics_pulse / toptica_pulse, Pulses.py:5-50) */
int zh_set_pulse_ex(zh_problem* p, int kind, double amp, const double* par,
                    const double* weights);

/* Piecewise cubic of ZH_PULSE_SPLINE: knots x (nknots, increasing) and coefficients
 * c (4, nknots-1) in scipy's CubicSpline.c layout, S(t) = sum_m c[m][i] (t-x_i)^(3-m)
 * on the interval i with x_i <= t (the last interval at the last knot). */
int zh_set_pulse_spline(zh_problem* p, int nknots, const double* x, const double* c);

/* The bias pulse of one instance (zh_set_pulse_ex's arguments; zh_set_pulse_ex sets
 * every instance's).  Replaces the per-run Bias.py of Zandpack/mpi/zand:113-124. */
int zh_instance_pulse(zh_problem* p, int instance, int kind, double amp, const double* par,
                      const double* weights);

/* Device-evaluated dissipator of the sigma equation, dissipator(t, sigma) =
 * eta(t) (sigma - sigma_eq), subtracted as D -= hbar * dissipator(t, sigma)
 * (Zandpack/mpi/zand:343, the Bias.dissipator hook; the form of
 * examples/Hubbard/Calculation/Bias.py:29-39).  kind: ZH_DISS_NONE (default),
 * ZH_DISS_CONST eta = eta0, ZH_DISS_STEP eta = eta0 / (1 + exp((t - t0) / width)).
 * sig_eq (NK, n, n) complex128 is required unless kind is ZH_DISS_NONE.  It enters
 * the attempts (zh_attempt*, zh_propagate); zh_rhs evaluates the in-process closure
 * f of the reference, which has no dissipator. */
int zh_set_dissipator(zh_problem* p, int kind, double eta0, double t0, double width,
                      const double* sig_eq);

/* Equilibrium density matrix by the contour sum of Zandpack/cmdtools/SCF:209-215
 * (the SCF tool's funci): rho = sum_i (i/2pi)(W_i G_i - conj(W_i) G_i^H),
 * G_i = (E_i - H - Sigma(E_i))^-1, one call per SCF iteration (its Jl.Parallel sum).
 *   contour  (nE, 4) real: Re E_i, Im E_i, Re W_i, Im W_i (already shifted by mu)
 *   evals    (NK, n) and evecs (NK, n, n) complex128: H = V diag(e) V^H per k
 *            (column o of V is eigenvector o)
 *   corb     (nc) orbitals that carry the lead self-energy, sigma_cc (NK, nE, nc, nc)
 *            complex128 its values Sigma(E_i) on them (FittedSelfEnergy.evaluate,
 *            summed over leads); nc = 0: no self-energy
 *   rho      (NK, n, n) complex128 out.
 * Stateless: allocates its device buffers on `device` and frees them. */
int zh_contour_density(int NK, int n, int nE, int nc, const double* contour,
                       const double* evals, const double* evecs, const int* corb,
                       const double* sigma_cc, double* rho, int device);

/* zh_contour_density without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving or SCF loops that must not call hipMalloc per iteration).  Same
 * arguments, results and errors as zh_contour_density (to rounding: the contour
 * points are summed in chunks that fit the scratch); its device buffers are slices of
 * a file-scope device array of ZH_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends, and it runs on the device's null stream.  It needs
 * 16 (4 n^2 + n nc + 3 P n nc + 4 P nc^2) bytes plus a few kB for P >= 1 contour
 * points per chunk (n = 1000, nc = 80: P = 16); larger n returns ZH_EINVAL before
 * any device work.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (from several threads) race and must be serialised by the caller. */
#define ZH_NOALLOC_SCRATCH_BYTES (128u << 20)
int zh_contour_density_noalloc(int NK, int n, int nE, int nc, const double* contour,
                               const double* evals, const double* evecs, const int* corb,
                               const double* sigma_cc, double* rho, int device);

/* Bias-window (nonequilibrium) density of Zandpack/cmdtools/SCF:236-253, 299-306
 * (G_l integrated over [mu_min - fact_kT kT_max, mu_max + fact_kT kT_max]):
 * rho = sum_i w_i (1/2pi) G(E_i)[:, c] M_i G(E_i)[:, c]^H on real nodes E_i, with
 * G = (E - H - Sigma(E))^-1 and M_i = sum_a (f_a(E_i) - f_min(E_i)) Gamma_a(E_i).
 *   nodes    (nE, 2) real: E_i, w_i (the caller's quadrature)
 *   evals, evecs, corb, sigma_cc as for zh_contour_density (sigma_cc at the nodes)
 *   gamma_w  (NK, nE, nc, nc) complex128: M_i on the coupled orbitals
 *   rho      (NK, n, n) complex128 out (Hermitian part).
 * Stateless like zh_contour_density. */
int zh_window_density(int NK, int n, int nE, int nc, const double* nodes, const double* evals,
                      const double* evecs, const int* corb, const double* sigma_cc,
                      const double* gamma_w, double* rho, int device);

/* Per-lead Pi of the resident state, as the stage-0 Pi the next attempt starts from
 * (the save_PI output of Zandpack/mpi/zand:376-378, 444-448): Q (NK, NL, n, n)
 * complex128 = (Pi_a + Pi_a^H)/hbar with hermitian = 1 (zand's Q_np, :90-110) or
 * Pi_a/hbar with hermitian = 0 (nozand's PI_np, mpi/nozand:120-140), where
 * Pi_a[o, o'] = sum over the modes m of lead a of psi_m[o] Ixi_m[o']. */
int zh_pi_get(zh_problem* p, int hermitian, double* Q);

/* Non-orthogonal basis (Zandpack/mpi/nozand): Sinv (NK, n, n) complex128 is S^-1
 * (nozand's Sm1 = S^(-0.5) S^(-0.5), :259-265); H0, xi, Ixi, sigma and psi are then
 * in the non-orthogonal basis.  S^-1 enters as nozand applies it: S^-1 H sigma in
 * the commutator (:456-461), Pi_a S^-1 in the sigma equation (:489-497), S^-1 H on
 * psi (:1075-1085), S^-1 on the GL_P xi constant (:1003) and on Omega xi
 * (:1208-1210); Omega, the currents and the error norm are unchanged in form.  An
 * explicit Ht (zh_rhs) is multiplied by S^-1 on the device.  Call after
 * zh_set_constants; NULL returns to the orthogonal form.  With the Hubbard term or
 * drive terms, S^-1 H(t, sigma) is formed on the device every stage.  Not combined
 * with zh_steady_state (ZH_EINVAL). */
int zh_set_overlap(zh_problem* p, const double* Sinv);

/* Error norm of the attempts.  (0, 0), the default, is zand's absolute norm
 * TE = sqrt(sum |err|^2) over sigma, Psi and Omega (Zandpack/mpi_funcs.py:247-252,
 * mpi/zand:390-396).  atol > 0 selects the scaled RMS norm
 * TE = sqrt(mean |err / (atol + rtol max(|y0|, |y1|))|^2) over the NK(n^2 + Mn + M^2)
 * entries, y0 and y1 the step's start and fifth-order end state, for runs specified
 * by rtol/atol (the reference's in-process drivers take atol/rtol, but their branch
 * evaluates to 1, TimedependentTransport.py:3226-3235); accept iff TE < eps (eps = 1
 * is the usual choice) and h_next = 0.9 h (eps/TE)^(1/5) as for the absolute norm. */
int zh_set_tolerance(zh_problem* p, double rtol, double atol);

/* Preconditioner of zh_steady_state: (chi+_i + Delta - H + B_i)^-1 per mode i, with
 * H = H(t0, sigma0) through its eigenpairs (evecs (NK,n,n) complex128, column k =
 * vector k; evals (NK,n) real) and, optionally, the lead self-energy term
 * B_i = sum_j xi_j Ixi_j^T i dGM_j / (chi-_j - chi+_i) of psinought's block
 * preconditioner (Zandpack/cmdtools/psinought:256-301) by the Woodbury identity on
 * the ncr orbitals corb: per group g of modes sharing chi+ (group: (NK, NL, NX, NOT),
 * reference mode order, ids < ngroups), K (ngroups, ncr, ncr) complex128 =
 * C_g (I + G_g C_g)^-1 with C_g the corb block of B and G_g that of
 * (chi+_g - H)^-1.  ncr = 0 leaves the self-energy out. */
typedef struct zh_precond {
  const double* evecs;
  const double* evals;
  int ncr;
  const int* corb;
  int ngroups;
  const int* group;
  const double* K;
} zh_precond;

/* Steady state of Zandpack/cmdtools/psinought (:180-449): Psi0 solving psinought's
 * real-linear system (Av3 / B3, :199-239) for the resident sigma taken as sigma0, the
 * bias at t0 and H(t0, sigma0) (Hubbard term included), and the closed-form Omega0
 * (:447-449); (sigma0, Psi0, Omega0) becomes the resident state.  BiCGSTAB over the
 * real inner product to ||b - A Psi|| <= tol ||b||, at most maxiter iterations,
 * right-preconditioned with pc (NULL: none; the bare operator's real spectrum
 * surrounds zero and rarely converges).  Returns ZH_ELIMIT when not converged (the
 * state then holds the last iterate); iters and rel_resid may be NULL and are set
 * in both cases. */
int zh_steady_state(zh_problem* p, double t0, double tol, int maxiter, const zh_precond* pc,
                    int* iters, double* rel_resid);

/* Resident state y0 = (sigma, Psi, Omega), reference layout, complex128.  A NULL
 * psi or omg puts zeros (cleared on the device, nothing uploaded): the usual start
 * of a propagation from sigma0 (Zandpack/mpi/zand:535-538). */
int zh_state_put(zh_problem* p, const double* sig, const double* psi, const double* omg);
int zh_state_get(zh_problem* p, double* sig, double* psi, double* omg);
int zh_sigma_get(zh_problem* p, double* sig);

/* One evaluation of the equations of motion, the reference's
 * f(t, sig, psi, omg, dH, delta, dH_given) (Zandpack/TimedependentTransport.py:2605-2660,
 * with the zand formulation Zandpack/mpi/zand:340-387, 746-809).  Ht is the full
 * Hamiltonian at this time (NK,n,n), delta the per-lead bias (NL).  Outputs are
 * the time derivatives; J (NL,NK) receives the lead currents of the input state
 * (Zandpack/mpi/zand:368-375) when non-NULL.  The resident state is untouched. */
int zh_rhs(zh_problem* p, const double* Ht, const double* delta, const double* sig,
           const double* psi, const double* omg, double* dsig, double* dpsi, double* domg,
           double* J);

/* One DOP54 attempt from the resident state (Zandpack/mpi/zand:301-408):
 * seven stages with device-evaluated H(t,sigma) and bias, the embedded error
 * norm and the step controller, all on the device.  When accepted the new
 * state replaces the resident one.  J (NL,NK) receives the stage-0 currents. */
int zh_attempt(zh_problem* p, zh_attempt_info* info, double* J);

/* The same attempt split in two, so attempts of several problems (each with its
 * own HIP stream) overlap on the device: launch enqueues it and returns; collect
 * waits for it, fills info (te, h_next, accepted) and J, and commits the state
 * when accepted.  One attempt per problem may be in flight.  No counterpart in
 * the reference (its ranks run one problem); used by the serving loop. */
int zh_attempt_launch(zh_problem* p, const zh_attempt_info* info);
int zh_attempt_collect(zh_problem* p, zh_attempt_info* info, double* J);
/* Non-blocking test of the attempt in flight: *ready = 1 when it has finished (the
 * next zh_attempt_collect returns at once), else 0.  For callers that collect each
 * problem as soon as its attempt is done (the serving loop). */
int zh_attempt_ready(zh_problem* p, int* ready);

/* One DOP54 attempt (Zandpack/mpi/zand:301-408) of every instance of the problem
 * together: info[g] (t0, h, eps, force_accept) in, (te, h_next, accepted) out for
 * instance g; J (NL, NK) the stage-0 currents of all k blocks.  Accepted instances
 * advance, the others keep their state; an instance whose next step size is not
 * finite is not committed (accepted = 0; h_next tells).  h = 0 with force_accept holds
 * an instance that has finished (its state stays bitwise; h_next is then not
 * meaningful).  launch / collect split it like zh_attempt_launch / _collect.  The
 * single-instance calls (zh_attempt, zh_propagate, zh_rhs, zh_steady_state) refuse a
 * problem with several instances. */
int zh_attempt_instances(zh_problem* p, zh_attempt_info* info, double* J);
int zh_launch_instances(zh_problem* p, const zh_attempt_info* info);
int zh_collect_instances(zh_problem* p, zh_attempt_info* info, double* J);

/* One attempt of each of count problems: every launch, then every collect, in one
 * call (the multi-instance step of the batched and serving configurations; the
 * reference runs one problem per process, Zandpack/mpi/zand:297-431 per problem).
 * infos[i] and J[i] (NL*NK doubles) are those of zh_attempt_collect for ps[i];
 * status[i] receives problem i's return code.  A failure of one problem leaves the
 * others running; the return value is the first failure's code (its message in
 * zh_last_error) or ZH_OK. */
int zh_batch_attempt(zh_problem* const* ps, int count, zh_attempt_info* infos, double* const* J,
                     int* status);

/* Tile autotuner: times a few compiled tile shapes of each GEMM of the right-hand
 * side on rejected attempts from the resident state (nothing is committed) and
 * keeps the fastest; choices are cached per (precision, size bucket) for the
 * process, so later problems of the same bucket reuse them (force = 1 re-times).
 * choice (optional) receives the 5 chosen candidate indices.  No counterpart in
 * the reference. */
int zh_autotune(zh_problem* p, int force, int* choice);

/* Set the 5 candidate indices directly (0..2 each), e.g. to reuse a tuning. */
int zh_set_tiles(zh_problem* p, const int* choice);

/* Device memory held by the problem (bytes), for admission by HBM footprint. */
int zh_problem_bytes(zh_problem* p, long* bytes);

/* Adaptive propagation loop (Zandpack/mpi/zand:297-431) from t0 while t < t1.
 * Writes up to cap accepted steps: times[i] (time after step i), J (cap,NL,NK)
 * stage-0 currents of step i, h_used[i].  *n_steps / *n_rejected are outputs.
 * Stops with ZH_ELIMIT after max_attempts attempts and ZH_ENONFINITE on a
 * non-finite error estimate. */
int zh_propagate(zh_problem* p, double t0, double t1, double h0, double eps, int time_update,
                 long max_attempts, long cap, double* times, double* J, double* h_used,
                 long* n_steps, long* n_rejected, double* h_final);

/* Fused Omega stage kernel of the mixed-precision attempts: 0 auto (default: used when
 * zh_autotune timed it faster than the unfused pair), 1 off, 2 on.  Fused, each stage's
 * Omega kernel also forms the Omega xi_c product of the Psi equation
 * (Zandpack/mpi/zand:771-778) from the k_s it has just stored (linearity over the
 * stages) instead of a separate product streaming k_s back; results agree with the
 * unfused path to rounding.  FP64 problems always run unfused. */
int zh_set_fusion(zh_problem* p, int mode);

/* Kernel of the Omega xi_c product of the Psi equation (Zandpack/mpi/zand:771-778):
 * 0 auto (default: the streaming kernel when zh_autotune timed it faster than the tiled
 * GEMM), 1 the tiled GEMM, 2 the streaming kernel (Omega and xi_c staged through LDS by
 * direct global-to-LDS loads, one lead's columns per tile).  The streaming kernel needs
 * every 16-column group of the coupled orbitals to belong to one lead, an even mode and
 * coupled-orbital count and at most 8 leads; where it does not apply, 2 falls back to the
 * GEMM (zh_omgxi_info tells which ran).  Both agree to rounding.  No counterpart in the
 * reference (a tensor contraction there). */
int zh_set_omgxi(zh_problem* p, int mode);

/* out[0] = 1 when the attempts run the streaming Omega xi_c kernel, out[1] = 1 when the
 * problem's layout allows it. */
int zh_omgxi_info(zh_problem* p, int* out);

/* How the Psi right-hand side runs with a row-compressed H (Zandpack/mpi/zand:746-778):
 * 0 auto (default: one pass when the rows of H reach at most 160 orbitals around each
 * 64-orbital tile, as tight-binding H do), 1 split (the dP xi_c sigma_c^T part as a
 * matrix-core product into the k slot, then the gather kernel), 2 one pass (stage-Psi
 * rows staged in LDS, that part formed beside the gather; ZH_EINVAL when H reaches
 * too far).  Both agree to rounding.  No counterpart in the reference. */
int zh_set_psi_kernel(zh_problem* p, int mode);
/* *fused = 1 when the attempts run the one-pass Psi kernel. */
int zh_psi_kernel_info(zh_problem* p, int* fused);

/* First same as last: 0 on (default), 1 off.  DOP54's seventh stage is evaluated at
 * (t0 + h, y_new); after an accepted attempt its k_7 = h f(t0 + h, y_new) is kept in the
 * k_1 slot and the next attempt (same state, step h') takes k_1 = (h'/h) k_7 instead of
 * evaluating its first stage (Zandpack/mpi/zand:303-387 evaluates all seven stages,
 * i == 0 included), when f is the same there: the stage parameters at its t0 (biases,
 * dissipator rate, drive factors; f depends on t only through them) equal those of that
 * seventh stage bit for bit, and nothing was changed in between (a new state, model,
 * tiles or kernels).  The currents of the first stage then come from the seventh
 * stage's y_new, the same state.  Decided per instance on the device; attempts of the
 * fused mixed-precision Omega kernel always evaluate it.  No counterpart in the
 * reference (a device schedule: same results to rounding). */
int zh_set_fsal(zh_problem* p, int mode);
/* *taken: instance attempts so far that took k_7 as their k_1; *on: 1 when the attempts
 * may (mode on and the kernels support it).  Either pointer may be null. */
int zh_fsal_info(zh_problem* p, long* taken, int* on);

/* Anti-Hermitian Omega: 0 auto (default), 1 off, 2 on.  dOmega_ij/dt (Zandpack/mpi/zand:
 * 784-809) maps Omega_ij -> -conj(Omega_ji) onto itself when every mode has
 * dGM_j Ixi_j = -conj(dGP_j xi_j) and Xpm_j = conj(Xpp_j) (the reference's Lorentzian
 * modes, Ixi = conj(xi) and dGP = dGM imaginary, TimedependentTransport.py:1032; modes
 * with dGP = dGM = 0, the Fermi poles), whatever psi and H are.  From an anti-Hermitian
 * Omega (zero, or psinought's Omega0) it stays so, and only its block upper triangle
 * (64 x 64 blocks of the device mode order) is computed and stored: half of the Omega
 * stage kernel's bytes and products.  Auto uses it when the constants have the symmetry
 * (to 1e-13 relative) and the Omega last given (zh_state_put, or zero) is anti-Hermitian
 * to 1e-13 of its largest entry; zh_rhs uses it per call under the same test of its
 * input.  zh_state_get and zh_rhs return Omega in full.  On (2) refuses constants or a
 * state without the symmetry (ZH_EINVAL).  No counterpart in the reference (which
 * computes every entry). */
int zh_set_antihermitian(zh_problem* p, int mode);

/* bit 0: Omega is held as its block upper triangle now; bit 1: the constants have the
 * symmetry; bit 2: the resident Omega is anti-Hermitian. */
int zh_antihermitian_info(zh_problem* p, int* out);

/* Use hipGraphs for zh_attempt / zh_propagate (1, default) or direct launches (0). */
int zh_set_graph(zh_problem* p, int enable);

/* Per-kernel-class HIP event timing on the problem's stream (1) or off (0, default).
 * While on, attempts run without graphs. */
int zh_set_profile(zh_problem* p, int enable);

/* Average device time per attempt and per kernel class, measured with HIP
 * events on the problem's stream since the last reset (ms).  names/ms get up
 * to cap entries; returns the number of entries. */
int zh_timing(zh_problem* p, int reset, int cap, char* names, double* ms, long* counts);

/* Building-block check: C = op(A) op(B) (complex128 in/out, host pointers) with
 * the engine's MFMA GEMM in FP64 (precision ZH_FP64) or FP32 (ZH_MIXED).
 * A is (M,K) or (K,M) when transA, B is (K,N) or (N,K) when transB. */
int zh_selftest_gemm(int precision, int M, int N, int K, int transA, int transB, const double* A,
                     const double* B, double* C);

#ifdef __cplusplus
}
#endif
#endif
