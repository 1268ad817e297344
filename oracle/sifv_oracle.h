/*
 * sifv_oracle.h - CPU restatement of the SemiImplicitFV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product path never links or calls it.
 *
 * Field arrays are passed as an array of SIFV_F_* indexed pointers
 * (NULL for absent fields) in the reference's padded x-fastest layout.
 */
#ifndef SIFV_ORACLE_H
#define SIFV_ORACLE_H

#include "../include/sifv_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NFIELDS 40

typedef struct or_solver or_solver;

or_solver* or_create(const sifv_mesh* mesh, const sifv_config* cfg, int riemann,
                     const sifv_eos* eos);
void or_destroy(or_solver* s);
size_t or_total_cells(const or_solver* s);
const char* or_last_error(void);

double or_explicit_step(or_solver* s, double* const* f, const sifv_config* cfg,
                        double targetDt);
double or_semi_step(or_solver* s, double* const* f, const sifv_config* cfg,
                    double targetDt, int* iters);
int or_jacobi(or_solver* s, const double* rho, const double* rhoc2, const double* rhs,
              double* p, double dt, double tol, int maxIter, double* hist);
int or_pcg(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* p, double dt,
           double tol, int maxIter);
int or_mgpcg(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* p, double dt,
             double tol, int maxIter);
int or_mg_precond(or_solver* s, const double* rho, const double* rhoc2, double dt, const double* r, double* z);
void or_set_pressure_solver(or_solver* s, int kind);
void or_fill_ghosts(or_solver* s, double* const* f, int varset);
void or_fill_scalar(or_solver* s, double* field);
void or_cons_to_prim(or_solver* s, double* const* f, const sifv_config* cfg);
double or_acoustic_dt(or_solver* s, double* const* f, const sifv_config* cfg,
                      double cfl, double maxDt);
double or_advective_dt(or_solver* s, double* const* f, double cfl, double maxDt);
/* Explicit RHS of the current state (after cons->prim and ghost fill have
 * been applied by the caller); rhs indexed like f (conservative ids). */
void or_compute_rhs(or_solver* s, double* const* f, const sifv_config* cfg,
                    double* const* rhs);
/* IGR entropic-pressure solve on the current primitive state. */
void or_igr(or_solver* s, double* const* f, const sifv_config* cfg);

/* Per-stencil reconstruction: order is one of SIFV_WENO[135] or SIFV_UPWIND[135], side 0 =
 * left-biased (v = 5/3 cells), side 1 = right-biased. */
double or_recon(int order, int side, const double* v, double eps);
/* Packed states: [rho,u0,u1,u2,p,sigma,gammaEff,piInfEff,alpha0..7] (16).
 * out: [mass,mom0,mom1,mom2,energy,alphaFlux0..7,pressureFlux,faceVelocity] (15). */
void or_riemann_flux(int type, const double* L, const double* R, int dir, int dim,
                     int includePressure, int useIGR, int nPhases, double* out);

#ifdef __cplusplus
}
#endif

#endif
