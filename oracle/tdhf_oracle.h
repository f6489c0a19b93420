/*
 * tdhf_oracle.h -- CPU restatement of mATTOHF's TDHF path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation.  Only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it; the product
 * library never links or calls it.
 *
 * Arrays follow the reference's Fortran layout: ubk(NG,NB,NK) complex128,
 * element (ig,ib,ik) at offset ig + NG*(ib + NB*ik) (0-based), stored as
 * interleaved (re,im) doubles.
 */
#ifndef TDHF_ORACLE_H
#define TDHF_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys {
  int ng, nk, nb, nbocc;
  double acell;
} orc_sys;

typedef struct orc_ctx orc_ctx;

/* tables and grids */
void   orc_grids(const orc_sys* s, double* rx, double* gx, double* kx);
long   orc_kvck_nsteps(const orc_sys* s);
double orc_k_vck(const orc_sys* s, double q);
void   orc_vcg(const orc_sys* s, double* vcg);
void   orc_vcgkk_full(const orc_sys* s, double* vcgkk);
void   orc_ion_potential(const orc_sys* s, int nion, const double* z, const double* alpha,
                         const double* rion, double* vion);
void   orc_init_wf(const orc_sys* s, double* ubk);
void   orc_gram_schmidt(const orc_sys* s, double* ubk);

/* FFT with FFTW's conventions (unnormalised; sign -1 forward, +1 backward) */
void   orc_fft(int n, int sign, const double* in, double* out);

/* context: owns vion and the Coulomb tables */
orc_ctx* orc_create(const orc_sys* s, const double* vion, const double* vcgkk_or_null);
void   orc_destroy(orc_ctx* c);
void   orc_set_threads(int n);

void   orc_dns(const orc_ctx* c, const double* ubk, double* dns);
void   orc_vh(const orc_ctx* c, const double* dns, double* vh);
void   orc_tpsi(const orc_ctx* c, const double* wbk, double A, double* out);
void   orc_vfpsi(const orc_ctx* c, const double* ubk, const double* wbk, double* out);
void   orc_hpsi(const orc_ctx* c, const double* ubk, const double* wbk, double A, double* out);
void   orc_ecomps(const orc_ctx* c, const double* ubk, double A, double* ecomps4);
double orc_etot(const orc_ctx* c, const double* ubk, double A);
double orc_je(const orc_ctx* c, const double* ubk, double A);
double orc_nele(const orc_ctx* c, const double* ubk);
void   orc_epsbk(const orc_ctx* c, const double* ubk, double* epsbk, double* epserr);

void   orc_rk4_step(const orc_ctx* c, double* ubk, double dt, double Aave, double A);
void   orc_rt_run(const orc_ctx* c, double* ubk, double dt, int nt, const double* A,
                  const double* Aave, double* etot, double* nele, double* jtot, double* init3);

void   orc_define_afield_lp(int nt, double dt, double omegac_ev, double tpulse_fs,
                            double e0_vnm, double phicep, int nenvelope,
                            double* time, double* A, double* Aave, double* E);
void   orc_define_afield_lr(int nt, double dt, double dAc, double* time, double* A,
                            double* Aave, double* E);

int    orc_heev(int m, double* a_colmajor, double* eigs);
void   orc_ks_krylov(const orc_ctx* c, double* ubk, int ncg);
void   orc_subspace_diag(const orc_ctx* c, double* ubk);
void   orc_gs_run(const orc_ctx* c, double* ubk, int nscf, int ncg, double* etot_hist,
                  double* epsbk, double* epserr);

void   orc_gs_run_diag(const orc_ctx* c, double* ubk, int nscf, int ncg, double* etot_hist,
                       double* epsbk, double* epserr, double* diag);
void   orc_make_dos(int nb, int nk, int nbocc, const double* epsbk, int ndos, double* epsdos,
                    double* dos, double* dosocc);
void   orc_fourier(int nt, double dt, const double* time, const double* J, double j0,
                   const double* E, const double* A, int nomega, double omegam, int window,
                   double tpulse, int nenvelope, double* out);
void   orc_te_step(const orc_ctx* c, double* ubk, double dt, double Aave, int variant);

double orc_adjoint_energy(const orc_ctx* c, const double* u0, double dt, int nt, const double* A,
                          double eref, double* dLdA, double* dLdu0);
double orc_objective_energy(const orc_ctx* c, const double* u0, double dt, int nt,
                            const double* A, double eref);
/* L = dt * sum_{it<nt} J(u_{it+1}, A[it])^2 (time-integrated current squared) */
double orc_adjoint_jsq(const orc_ctx* c, const double* u0, double dt, int nt, const double* A,
                       double* dLdA, double* dLdu0);
double orc_objective_jsq(const orc_ctx* c, const double* u0, double dt, int nt, const double* A);

#ifdef __cplusplus
}
#endif
#endif
