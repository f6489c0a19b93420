/*
 * tdhf.h -- C-ABI of the MI355X (gfx950) TDHF library, libtdhf.so.
 *
 * Drop-in boundary for mATTOHF's GS_calculation / RT_calculation call path.
 * The reference has no FFI: its path is plain Fortran subroutines over
 * module-global state (src/Global_variables.f90:5-57).  Every entry point
 * below names the reference routine it replaces.  A Fortran caller binds them
 * with iso_c_binding (mattohf_amd/fortran/tdhf_capi.f90).
 *
 * Conventions
 *  - Orbital arrays are HOST arrays in the reference's Fortran layout
 *    ubk(NG,NB,NK) complex(8): element (ig,ib,ik) (0-based) at offset
 *    ig + NG*(ib + NB*ik), stored as interleaved (re,im) doubles; a
 *    complex(c_double_complex) Fortran array can be passed directly.
 *  - The caller owns all host buffers; the context owns device memory,
 *    streams and graphs.  Calls are synchronous at return.
 *  - Return value 0 = success, negative = error (TDHF_E*); the message is
 *    available from tdhf_last_error().  No exception crosses the ABI.
 *  - A context (or batch) is used by one host thread at a time; several may
 *    coexist and be used from different threads at once (each owns its own
 *    device buffers and stream; the library's shared records are guarded).
 *  - Everything is FP64 / complex128.
 */
#ifndef TDHF_H
#define TDHF_H

#ifdef __cplusplus
extern "C" {
#endif

#define TDHF_OK 0
#define TDHF_EINVAL (-1)  /* invalid argument or unsupported size */
#define TDHF_ENOMEM (-2)  /* device allocation failed */
#define TDHF_EDEVICE (-3) /* HIP runtime error or no gfx950 device */
#define TDHF_ENUMERIC (-4) /* non-finite values detected */

/* System description; mirrors src/Global_variables.f90:29-44 (NG, NK, NB,
 * NBocc, acell) and the RT time step dt (:13).  Occupations are 2/NK for the
 * first NBocc bands and 0 otherwise (src/Initialization.f90:176-186).
 * Supported NG: 16, 32, 64, 128. */
typedef struct tdhf_params {
  int ng;
  int nk;
  int nb;
  int nbocc;
  double acell;
  double dt;
} tdhf_params;

typedef struct tdhf_ctx tdhf_ctx;

/* Creates a context for one system.  Replaces the table set-up of
 * Initialization (src/Initialization.f90:86-146,176-186).
 *   vion  : NG doubles, the ion potential (required; see tdhf_ion_potential)
 *   vcg   : NG doubles vCG(G) with vCG(G=0)=0, or NULL to generate on device
 *   vcgkk : NG*NK*NK doubles in the reference layout vCGkk(ig, ikk) with
 *           ikk = ikr*NK + ikc, or NULL to generate the k_vCk table on the
 *           device (src/Functions.f90:33-52).  A supplied table must depend
 *           on ikr-ikc only; the library keeps the NG*(2NK-1) distinct rows. */
int tdhf_create(const tdhf_params* p, const double* vion, const double* vcg,
                const double* vcgkk, tdhf_ctx** out);

/* Releases a context and all of its device memory. */
int tdhf_destroy(tdhf_ctx* ctx);

/* Message of the last failed call on ctx (or of the last failed
 * tdhf_create / context-free call when ctx is NULL).  Never NULL. */
const char* tdhf_last_error(const tdhf_ctx* ctx);

/* Ion potential on the device.  Replaces Prep_periodic_ion_potential
 * (src/Prep_periodic_ion_potential.f90:10-20): images |n| <= 4*NK of
 * -Z/sqrt(d^2+alpha^2), then shifted so that max(vion) = 0.
 *   z, alpha, rion : nion values each (per ion; rion in units of acell). */
int tdhf_ion_potential(const tdhf_params* p, int nion, const double* z, const double* alpha,
                       const double* rion, double* vion);

/* Coulomb tables as used by a context, for inspection: vcg (NG) and the
 * distinct exchange-kernel rows vq(ig, d) for d = ikr-ikc+NK-1 in [0,2NK-1),
 * i.e. NG*(2NK-1) doubles, column d at offset NG*d.  Either may be NULL.
 * Replaces the tables of src/Initialization.f90:116-146. */
int tdhf_get_tables(tdhf_ctx* ctx, double* vcg, double* vq);

/* H[u] w.  Replaces ubkwbk_hwbk(ubk,wbk,A) (src/Functions.f90:187-209):
 * kinetic (src/Functions.f90:212-234) + (vH[rho[u]] + vion) w + Fock
 * exchange of w with the occupied bands of u (src/Functions.f90:237-261).
 *   u, w, hw : NG*NB*NK complex each (hw may alias neither u nor w). */
int tdhf_hpsi(tdhf_ctx* ctx, const double* u, const double* w, double A, double* hw);

/* Observables of u at vector potential A.  Replaces ubk_Ecomps / ubk_Etot
 * (src/Functions.f90:74-91,265-286), ubk_Je (:289-318) and the electron
 * number of src/RT_calculation.f90:43-45.
 *   ecomps : (Ekin, Eion, EH, Ex), may be NULL; other outputs may be NULL. */
int tdhf_observables(tdhf_ctx* ctx, const double* u, double A, double* etot, double* ecomps,
                     double* nele, double* je);

/* One RK4 step in place.  Replaces RT_RK4_ubk(ubk,Aave,A)
 * (src/RT_calculation.f90:154-173): stages at A(it-1)=2*Aave-A, Aave, Aave, A. */
int tdhf_rk4_step(tdhf_ctx* ctx, double* u, double Aave, double A);

/* tdhf_rk4_step for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  u is a DEVICE pointer to the state
 * (NG*NB*NK complex, reference layout), advanced in place by one RK4 step at
 * A(it-1) = 2*Aave - A, Aave, Aave, A; the four stage launches are enqueued on
 * stream (a hipStream_t; NULL = the context's stream) and the call returns
 * without waiting.  The stage buffers live in a fixed device array of the
 * library (6 MiB per GPU, held until the process ends): NG*NB*NK must not
 * exceed 131072 (TDHF_EINVAL otherwise).  The call allocates nothing; the
 * context (tables) must have been created before.  All calls on one GPU share
 * that array, so calls whose launches overlap on one GPU (other streams,
 * other host threads) race. */
int tdhf_rk4_step_noalloc(tdhf_ctx* ctx, double* u, double Aave, double A, void* stream);

/* nt RK4 steps with the observables after every step.  Replaces the loop of
 * RT_calculation (src/RT_calculation.f90:29-60).  u is propagated in place.
 *   A, Aave          : nt values (Define_Afield output, src/RT_calculation.f90
 *                      :175-244); step it (0-based) uses A[it-1] (0 for it=0),
 *                      Aave[it], A[it]
 *   etot, nele, jtot : nt values each, observables of the state after step it
 *                      evaluated at A[it]; may be NULL
 *   init3            : (Etot, N, J) of the input state at A=0, may be NULL. */
int tdhf_rt_run(tdhf_ctx* ctx, double* u, int nt, const double* A, const double* Aave,
                double* etot, double* nele, double* jtot, double* init3);

/* Device-resident variant used for timing: tdhf_state_upload copies u to the
 * device once; tdhf_rt_propagate then runs nt steps on the resident state
 * (observables as in tdhf_rt_run, written to host arrays at the end);
 * tdhf_state_download copies the state back. */
int tdhf_state_upload(tdhf_ctx* ctx, const double* u);
int tdhf_state_download(tdhf_ctx* ctx, double* u);
int tdhf_rt_propagate(tdhf_ctx* ctx, int nt, const double* A, const double* Aave, double* etot,
                      double* nele, double* jtot, double* init3);

/* The same loop split into calls, for callers that keep the field on the
 * device and advance the resident state in pieces (the time loop of
 * RT_calculation, src/RT_calculation.f90:38-60, with its field arrays from
 * Define_Afield, :175-244, uploaded once).
 *   tdhf_rt_begin   : uploads A, Aave (nt values each, as tdhf_rt_run) and
 *                     sets the step counter to 0; the state is the one last
 *                     uploaded (tdhf_state_upload) or propagated
 *   tdhf_rt_advance : the next n steps (counting from the begin), in place
 *                     on the device state; fails if they exceed nt
 *   tdhf_rt_end     : observables of every step advanced since the begin
 *                     (done values each, as tdhf_rt_run; NULL ok) and init3;
 *                     closes the run
 * tdhf_rt_propagate(nt, ...) is begin(nt) + advance(nt) + end.  The state can
 * be downloaded between advances (tdhf_state_download). */
int tdhf_rt_begin(tdhf_ctx* ctx, int nt, const double* A, const double* Aave);
int tdhf_rt_advance(tdhf_ctx* ctx, int n);
int tdhf_rt_end(tdhf_ctx* ctx, double* etot, double* nele, double* jtot, double* init3);

/* Ground state.  Replaces GS_calculation (src/GS_calculation.f90:26-57) with
 * KS_ubk2_SVD (:124-194), Subspace_diag (:197-237) and Gram_Schmidt
 * (src/mATTOHF.f90:134-157).  Runs nscf iterations (or stops early once
 * |E(iter)-E(iter-1)| < tol when tol > 0).
 *   etot_hist : nscf+1 values (entry 0 = input orbitals), may be NULL
 *   epsbk, epserr : NB*NK band energies / residuals of the final orbitals
 *                   (ubk_epsbk, ubk_epserrorbk, src/Functions.f90:321-358),
 *                   may be NULL
 *   niter     : iterations performed, may be NULL. */
int tdhf_gs_run(tdhf_ctx* ctx, double* u, int nscf, int ncg, double tol, double* etot_hist,
                double* epsbk, double* epserr, int* niter);

/* Gram_Schmidt (src/mATTOHF.f90:134-157) on the device, in place. */
int tdhf_gram_schmidt(tdhf_ctx* ctx, double* u);

/* Per-kernel timing for measurement, with HIP events on the stream the
 * dominant kernel (the fused H-apply/RK4-stage kernel) is launched on.
 * enable = 1: every launch is bracketed by an event pair (graph replay off);
 * enable = 2: the RK4 steps of each tdhf_rt_propagate call (graph-replayed
 * chunks and remainder) are bracketed by one event pair and count 4 launches
 * per step (no per-launch event overhead); 0: off.  tdhf_kernel_stats returns the number of timed
 * launches and their summed duration in milliseconds since the last reset. */
int tdhf_set_kernel_timing(tdhf_ctx* ctx, int enable);
int tdhf_kernel_stats(tdhf_ctx* ctx, long* launches, double* total_ms);

/* RT_TE_ubk (variant 0), RT_TEPC_ubk (1), RT_TEPC2_ubk (2),
 * src/RT_calculation.f90:71-151: one step of the 4th-order Taylor
 * propagators with H frozen at the step start (TE), with a predictor-
 * corrector average of H (TEPC) or H at the predictor midpoint (TEPC2), at
 * field Aave; u in/out (NG*NB*NK complex). */
int tdhf_te_step(tdhf_ctx* ctx, double* u, double Aave, int variant);

/* GS_calculation (src/GS_calculation.f90:2-121) with the per-iteration
 * columns of <sys_name>_GSinfo.out: diag[5*(iter-1) + ...] = mean |change of
 * the mixed density dnsscf| (mixing 1/2), mean and max epserrorbk,
 * fundamental gap, smallest same-k gap (DBL_MAX when NB == NBocc, where the
 * reference's minval runs over an empty set); nscf iterations, no early
 * stop.  etot_hist: nscf+1; epsbk/epserr: NB*NK of the final state. */
int tdhf_gs_run_diag(tdhf_ctx* ctx, double* u, int nscf, int ncg, double* etot_hist,
                     double* epsbk, double* epserr, double* diag);

/* Make_DoS with GS_calculation's energy grid (src/GS_calculation.f90:93-97,
 * 239-265): epsdos[i] = emin + (emax-emin) i/ndos with emin/emax the band
 * extremes -/+ 10 % of their range, Gaussian width 5 (epsdos[1]-epsdos[0]);
 * dos weighted by the k-point's largest occupation, dosocc by the band's
 * (2/NK for ib < NBocc).  epsbk: NB*NK. */
int tdhf_make_dos(int nb, int nk, int nbocc, const double* epsbk, int ndos, double* epsdos,
                  double* dos, double* dosocc);

/* FourierTr (window 0) / WindowFourierTr (window 1), src/RT_calculation.f90
 * :246-351: out[7*i + ...] = omega_i = (i+1)/nomega*omegam, Re JF, Im JF,
 * Re EF, Im EF, Re AF, Im AF with XF = sum_t exp(i omega t) filter(t) X(t) dt
 * (X = J - j0, E, A); filter 0 = 1 - 3 (t/t_N)^2 + 2 (t/t_N)^3, filter 1 =
 * sin(pi t/tpulse)^nenvelope on the first int(min(tpulse/dt, nt)) steps.
 * time, J, E, A: nt values (time as Define_Afield: dt*(it-1)). */
int tdhf_fourier(int nt, double dt, const double* time, const double* J, double j0,
                 const double* E, const double* A, int nomega, double omegam, int window,
                 double tpulse, int nenvelope, double* out);

/* FP64 FMA throughput of the device (no counterpart in the reference: a
 * measurement probe for the benchmark).  Runs iters rounds of 8 independent
 * fused multiply-adds per thread on 1024 x 256 threads, once to warm up and
 * three times timed; *tflops = 2 * 8 * iters * 262144 / (fastest time),
 * *ms = that time (may be NULL). */
int tdhf_fp64_fma_probe(int iters, double* tflops, double* ms);

/* ------------------------------------------------------------------------
 * Differentiation.  No counterpart in the reference: a hand-written adjoint
 * of the RK4 propagation (RT_RK4_ubk, src/RT_calculation.f90:154-173, with
 * H = ubkwbk_hwbk, src/Functions.f90:187-209) for the objective
 *     L(A) = Etot(u_N, A[N-1]) - eref        (Etot as ubk_Etot, :74-91)
 * after nt steps driven by the per-step field A[0..nt-1] (Aave is derived
 * from A as Define_Afield does, src/RT_calculation.f90:225-230).
 * tdhf_fwd runs the forward pass and keeps either every RK4 stage state or,
 * when those exceed mem_budget bytes (<= 0: a quarter of free device
 * memory), a checkpoint every S = ceil(sqrt(nt)) steps (or ~sqrt(nt/4), the
 * least-memory spacing, if that does not fit) with segment recomputation in
 * tdhf_bwd.  Requires NB == NBocc.
 *   u     : initial orbitals in, final orbitals out
 *   L     : objective; etot/nele/jtot: nt observables as tdhf_rt_run (NULL ok)
 * tdhf_bwd returns dL/dA[0..nt-1] and, if dLdu0 != NULL, dL/du0 as a
 * Euclidean cotangent (dL = Re sum conj(dLdu0) du0), NG*NB*NK complex. */
int tdhf_fwd(tdhf_ctx* ctx, double* u, int nt, const double* A, double eref, double mem_budget,
             double* L, double* etot, double* nele, double* jtot);
int tdhf_bwd(tdhf_ctx* ctx, double* dLdA, double* dLdu0);

/* Objectives of the forward/backward pair. */
#define TDHF_OBJ_ENERGY 0 /* L = Etot(u_N, A[N-1]) - eref                          */
#define TDHF_OBJ_JSQ 1    /* L = dt * sum_{n=1..N} J_n^2, J_n = ubk_Je(u_n, A[n-1])  */

/* The same pair for ntraj trajectories of the context's cell at once (own
 * initial orbitals and field each; same NK, tables and ion potential), one
 * launch per RK4 stage for all of them (adjoint stages likewise), with either
 * objective.  The memory budget covers the stage states of all trajectories.
 * At NG = 128 the backward call evaluates the reverse stages by pair tiles
 * (one pair of transforms per unordered pair of rows) and allocates their
 * partial sums outside the budget, ntraj * NK*NB * ceil(NK*NB/32) * NG * 16
 * bytes (2.1 GB for 32 trajectories at NK = 128, NB = 8); without that
 * memory it evaluates them row-wise (the same results to rounding).
 *   u      : ntraj states concatenated (NG*NB*NK complex each), in: initial,
 *            out: final
 *   A      : ntraj*nt, trajectory t at A + nt*t
 *   eref   : ntraj reference energies (TDHF_OBJ_ENERGY; NULL = 0)
 *   L      : ntraj objectives; etot/nele/jtot: ntraj*nt (NULL ok)
 * tdhf_bwd_batch: dLdA ntraj*nt; dLdu0 (NULL ok) ntraj states.
 * tdhf_fwd / tdhf_bwd are this pair with ntraj = 1 and TDHF_OBJ_ENERGY. */
int tdhf_fwd_batch(tdhf_ctx* ctx, int ntraj, double* u, int nt, const double* A, int objective,
                   const double* eref, double mem_budget, double* L, double* etot, double* nele,
                   double* jtot);
int tdhf_bwd_batch(tdhf_ctx* ctx, double* dLdA, double* dLdu0);
/* How the last tdhf_fwd stored its states: full_tape = 1 if every stage state
 * is kept, else segment = checkpoint spacing S. */
int tdhf_adjoint_info(tdhf_ctx* ctx, int* full_tape, int* segment);
/* Progress of the checkpointed adjoint calls (tdhf_fwd_batch, tdhf_bwd_batch
 * and their single-pulse forms): with fn set, the call waits for each
 * checkpoint segment to finish on the device and then calls
 * fn(user, phase, done, total) on the calling thread (phase 0: forward
 * segments, 1: backward segments, each counted 1..total).  fn = NULL (the
 * default) turns it off.  A full-tape run reports nothing.  The reference's
 * counterpart is the RT wall time it prints every 1000 steps
 * (src/RT_calculation.f90:56-58). */
typedef void (*tdhf_progress_fn)(void* user, int phase, long done, long total);
int tdhf_set_progress(tdhf_ctx* ctx, tdhf_progress_fn fn, void* user);

/* ------------------------------------------------------------------------
 * Batched trajectories.  No counterpart in the reference (which runs one
 * trajectory per process): many independent RT runs of one cell (NG, NB,
 * NBocc, acell, dt shared; NK, ion potential and field per trajectory)
 * advance together, one launch per RK4 stage for all of them, the steps
 * replayed from a captured hipGraph.  Each trajectory is exactly the run
 * tdhf_rt_run would make with its own context.
 *   nk[t]   : k-points of trajectory t (p->nk is ignored)
 *   vion    : NG*ntraj doubles, trajectory t at vion + NG*t
 *   u       : all orbitals concatenated; trajectory t is (NG,NB,nk[t])
 *             complex starting after those of trajectories 0..t-1
 *   A, Aave : ntraj*nt doubles, trajectory t at offset nt*t
 *   etot, nele, jtot : ntraj*nt (may be NULL); init3 : ntraj*3 (may be NULL)
 * ------------------------------------------------------------------------ */
typedef struct tdhf_batch tdhf_batch;

int tdhf_batch_create(const tdhf_params* p, int ntraj, const int* nk, const double* vion,
                      tdhf_batch** out);
int tdhf_batch_destroy(tdhf_batch* b);
const char* tdhf_batch_last_error(const tdhf_batch* b);
/* traj >= 0: one trajectory's orbitals; traj < 0: all, concatenated. */
int tdhf_batch_upload(tdhf_batch* b, int traj, const double* u);
int tdhf_batch_download(tdhf_batch* b, int traj, double* u);
/* nt RK4 steps of every trajectory with the observables after each step. */
int tdhf_batch_propagate(tdhf_batch* b, int nt, const double* A, const double* Aave,
                         double* etot, double* nele, double* jtot, double* init3);
/* create + upload + propagate + download + destroy in one call. */
int tdhf_batch_run(const tdhf_params* p, int ntraj, const int* nk, const double* vion,
                   double* u, int nt, const double* A, const double* Aave, double* etot,
                   double* nele, double* jtot, double* init3);
/* Launch shape: waves per workgroup of the fused kernel, or with pair tiles of
 * the per-row epilogue launch (1, 2, 4 or 8, at least NG/64; default the
 * smallest).
 * tdhf_batch_autotune times nt_trial steps for each candidate on a copy of the
 * state (the batch state is unchanged), keeps the fastest and remembers it for
 * batches of the same shape created later in the process; best_nw receives
 * its waves per workgroup (tdhf_batch_get_schedule its row schedule). */
int tdhf_batch_set_waves(tdhf_batch* b, int nw);
int tdhf_batch_get_waves(const tdhf_batch* b);
int tdhf_batch_autotune(tdhf_batch* b, int nt_trial, const double* A, const double* Aave,
                        int* best_nw);
/* Row schedule of the batched launch (which workgroup runs which trajectory's
 * row, and so on which XCD): 0 = trajectories bin-packed over the XCDs (a
 * trajectory's rows on one XCD, contiguous; the default), 1 = the same with
 * the rows of an XCD's trajectories interleaved, 2 = no XCD pinning.  The
 * results do not depend on it.  tdhf_batch_autotune times every (exchange
 * evaluation, schedule, waves per workgroup) combination and keeps the
 * fastest. */
int tdhf_batch_set_schedule(tdhf_batch* b, int sched);
int tdhf_batch_get_schedule(const tdhf_batch* b);
/* Exchange evaluation of the batched stages: 0 = row-wise (every ordered pair
 * of rows (k,i), (q,j) transformed by the fused kernel of row (k,i), as
 * ubkwbk_vFwbk, src/Functions.f90:247-257), 1, 2, 3 = pair tiles of 64, 32
 * or 16 rows: since the stage acts on its own state, the convolutions of
 * (k,i),(q,j) and (q,j),(k,i) are conjugates up to the Nyquist component, so
 * one transform pair per unordered pair feeds both rows (needs NG = 64,
 * NB == NBocc and fewer than 2^22 rows per trajectory; TDHF_EINVAL
 * otherwise); a second launch per stage adds each row's partial sums and
 * applies the rest of H and the stage update.  Smaller tiles make more
 * workgroups and more partial sums per row.  Results agree to rounding.  tdhf_batch_autotune includes all four. */
int tdhf_batch_set_exchange(tdhf_batch* b, int mode);
int tdhf_batch_get_exchange(const tdhf_batch* b);
int tdhf_batch_set_kernel_timing(tdhf_batch* b, int enable);
int tdhf_batch_kernel_stats(tdhf_batch* b, long* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* TDHF_H */
