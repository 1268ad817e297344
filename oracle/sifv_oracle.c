/*
 * sifv_oracle.c - plain-C restatement of the SemiImplicitFV hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU checker for the HIP library.  It is
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, never by the product path.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the reference repository root).  Arithmetic is written in the
 * reference's evaluation order and built with -ffp-contract=off, so the
 * restatement agrees with the compiled reference to rounding (~1e-15).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   recon, riemann flux, cons->prim, ghost fill, dt, explicit step, RHS,
 *   Jacobi, semi-implicit step, IGR and mixture EOS are each checked against
 *   the compiled reference (oracle/_ref) and against committed fixtures in
 *   tests/golden/, and by closed-form properties (polynomial exactness of
 *   the upwind/WENO weights, flux consistency F(U,U)=F(U), conservation,
 *   Jacobi fixed point = direct solve, published 1D advection error norms).
 */
#include "sifv_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NF OR_NFIELDS
#define MAXPH SIFV_MAX_PHASES

static char g_err[256];
const char* or_last_error(void) { return g_err; }

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* ------------------------------------------------------------------ */
/* mesh: include/RectilinearMesh.hpp:60-105, src/RectilinearMesh.cpp:14-114 */
/* ------------------------------------------------------------------ */
typedef struct {
    int dim, nx, ny, nz, ng, ngx, ngy, ngz, nxT, nyT, nzT;
    size_t total;
    double *xe, *ye, *ze; /* extended node arrays */
    int bc[6];
    int wrap[6];    /* face wraps periodically (Periodic BC or MPI periods) */
    int mpiwrap[6]; /* wrap comes from MPI periods (halo self exchange) */
} or_mesh;

/* buildExtendedNodes, src/RectilinearMesh.cpp:91-114 */
static double* build_ext(const double* phys, int n, int ng) {
    int next = n + 2 * ng + 1;
    double* e = (double*)malloc(sizeof(double) * next);
    for (int i = 0; i <= n; ++i) e[ng + i] = phys[i];
    for (int g = 1; g <= ng; ++g) {
        int m = (g - 1 < n - 1) ? g - 1 : n - 1;
        double w = phys[m + 1] - phys[m];
        e[ng - g] = e[ng - g + 1] - w;
    }
    for (int g = 1; g <= ng; ++g) {
        int m = (n - g > 0) ? n - g : 0;
        double w = phys[m + 1] - phys[m];
        e[ng + n + g] = e[ng + n + g - 1] + w;
    }
    return e;
}

static inline size_t IDX(const or_mesh* m, int i, int j, int k) {
    return (size_t)((i + m->ngx) + m->nxT * ((j + m->ngy) + m->nyT * (k + m->ngz)));
}
static inline double DX(const or_mesh* m, int i) { return m->xe[i + m->ngx + 1] - m->xe[i + m->ngx]; }
static inline double DY(const or_mesh* m, int j) { return m->ye[j + m->ngy + 1] - m->ye[j + m->ngy]; }
static inline double DZ(const or_mesh* m, int k) { return m->ze[k + m->ngz + 1] - m->ze[k + m->ngz]; }
static inline double VOL(const or_mesh* m, int i, int j, int k) { return DX(m, i) * DY(m, j) * DZ(m, k); }

/* ------------------------------------------------------------------ */
/* solver object                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    double rho, u[3], p, sigma, gammaEff, piInfEff, alpha[MAXPH];
} prim_t; /* PrimitiveState, include/State.hpp:19-30 (T unused on the path) */

typedef struct {
    double mass, mom[3], energy, alphaFlux[MAXPH], pressureFlux, faceVelocity;
} flux_t; /* RiemannFlux, include/RiemannSolver.hpp:26-36 */

struct or_solver {
    or_mesh m;
    int riemann;
    int eosKind;
    double gamma, pInf;
    /* scratch: rhs per conservative id, RK backups, semi-implicit arrays */
    double* rhs[NF];
    double* bak[NF];
    double* rhsP;
    double* divU;
    double* rhoUStar[3];
    double* rhoEstar;
    double* pAdv;
    double* rhoc2;
    double* divUstar;
    double* pRhs;
    double* pressure_;
    double* pNew;
    double* grad; /* 9 per cell */
    prim_t* faceL;
    prim_t* faceR;
    int psolver;  /* pressure solver of the semi-implicit step: 0 Jacobi, 1 PCG, 2 multigrid PCG */
    double* cg[6]; /* PCG vectors r, z, p, q, d, s (lazily allocated) */
    int mgLevels;  /* coarse levels of the multigrid preconditioner (0: not built) */
    struct or_mglev* mg;
};

size_t or_total_cells(const or_solver* s) { return s->m.total; }

static double* zalloc(size_t n) { return (double*)calloc(n, sizeof(double)); }

/* One coarse level of the multigrid preconditioner: n cells per axis,
 * compact x-fastest arrays (no ghosts).  s: mass term V/c summed over the
 * children; w[d]: weight of the cell's low face along d (the wrap face at
 * index 0 of a periodic axis, else 0 there); d: diagonal; b, x, t: right-hand
 * side, iterate and output of the V-cycle on this level. */
struct or_mglev {
    int n[3];
    size_t cells;
    double *s, *w[3], *d, *b, *x, *t;
};

static void mg_free(struct or_solver* s) {
    for (int l = 0; l < s->mgLevels; ++l) {
        struct or_mglev* L = &s->mg[l];
        free(L->s);
        for (int d = 0; d < 3; ++d) free(L->w[d]);
        free(L->d);
        free(L->b);
        free(L->x);
        free(L->t);
    }
    free(s->mg);
    s->mg = NULL;
    s->mgLevels = 0;
}

or_solver* or_create(const sifv_mesh* mesh, const sifv_config* cfg, int riemann,
                     const sifv_eos* eos) {
    static const double unit[2] = {0.0, 1.0};
    if (mesh->dim < 1 || mesh->dim > 3 || mesh->nx < 1) {
        snprintf(g_err, sizeof g_err, "bad mesh");
        return NULL;
    }
    or_solver* s = (or_solver*)calloc(1, sizeof(or_solver));
    or_mesh* m = &s->m;
    m->dim = mesh->dim;
    m->ng = mesh->nGhost;
    m->nx = mesh->nx;
    m->ny = mesh->dim >= 2 ? mesh->ny : 1;
    m->nz = mesh->dim >= 3 ? mesh->nz : 1;
    m->ngx = m->ng;
    m->ngy = m->dim >= 2 ? m->ng : 0;
    m->ngz = m->dim >= 3 ? m->ng : 0;
    m->nxT = m->nx + 2 * m->ngx;
    m->nyT = m->ny + 2 * m->ngy;
    m->nzT = m->nz + 2 * m->ngz;
    m->total = (size_t)m->nxT * m->nyT * m->nzT;
    m->xe = build_ext(mesh->xNodes, m->nx, m->ngx);
    m->ye = build_ext((m->dim >= 2 && mesh->yNodes) ? mesh->yNodes : unit, m->ny, m->ngy);
    m->ze = build_ext((m->dim >= 3 && mesh->zNodes) ? mesh->zNodes : unit, m->nz, m->ngz);
    for (int f = 0; f < 6; ++f) {
        int d = f / 2;
        int periodic = mesh->periods[d] != 0;
        /* Runtime::setBoundaryCondition is a no-op on periodic (non-physical)
         * faces (src/Runtime.cpp:123-127); the halo self-exchange wraps. */
        m->bc[f] = periodic ? SIFV_BC_OUTFLOW : mesh->bc[f];
        m->wrap[f] = periodic || mesh->bc[f] == SIFV_BC_PERIODIC;
        m->mpiwrap[f] = periodic;
    }
    s->riemann = riemann;
    s->eosKind = eos->kind;
    s->gamma = eos->gamma;
    s->pInf = eos->kind == SIFV_EOS_STIFFENED ? eos->pInf : 0.0;
    size_t n = m->total;
    for (int f = 0; f < NF; ++f) {
        s->rhs[f] = zalloc(n);
        s->bak[f] = zalloc(n);
    }
    s->rhsP = zalloc(n);
    s->divU = zalloc(n);
    for (int d = 0; d < 3; ++d) s->rhoUStar[d] = zalloc(n);
    s->rhoEstar = zalloc(n);
    s->pAdv = zalloc(n);
    s->rhoc2 = zalloc(n);
    s->divUstar = zalloc(n);
    s->pRhs = zalloc(n);
    s->pressure_ = zalloc(n);
    s->pNew = zalloc(n);
    s->grad = zalloc(9 * n);
    size_t nfx = (size_t)(m->nx + 1) * m->ny * m->nz;
    size_t nfy = (size_t)m->nx * (m->ny + 1) * m->nz;
    size_t nfz = (size_t)m->nx * m->ny * (m->nz + 1);
    size_t nf = nfx;
    if (nfy > nf) nf = nfy;
    if (nfz > nf) nf = nfz;
    s->faceL = (prim_t*)calloc(nf, sizeof(prim_t));
    s->faceR = (prim_t*)calloc(nf, sizeof(prim_t));
    (void)cfg;
    return s;
}

void or_destroy(or_solver* s) {
    if (!s) return;
    free(s->m.xe);
    free(s->m.ye);
    free(s->m.ze);
    for (int f = 0; f < NF; ++f) {
        free(s->rhs[f]);
        free(s->bak[f]);
    }
    free(s->rhsP);
    free(s->divU);
    for (int d = 0; d < 3; ++d) free(s->rhoUStar[d]);
    free(s->rhoEstar);
    free(s->pAdv);
    free(s->rhoc2);
    free(s->divUstar);
    free(s->pRhs);
    free(s->pressure_);
    free(s->pNew);
    for (int q = 0; q < 6; ++q) free(s->cg[q]);
    mg_free(s);
    free(s->grad);
    free(s->faceL);
    free(s->faceR);
    free(s);
}

/* ------------------------------------------------------------------ */
/* ghost fill                                                           */
/* ------------------------------------------------------------------ */
/* copyCell_P / copyCell_C / copyCell, src/SolutionState.cpp:79-147.  For
 * faces wrapped by the MPI self exchange the copied set is the halo pack's
 * (src/HaloExchange.cpp:206-262): identical except that the CONS pack
 * carries no sigma. */
static void copy_cell(const or_mesh* m, double* const* f, int varset, size_t dst, size_t src,
                      double sU, double sV, double sW, int exch) {
    const int vel[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    const int mom[3] = {SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW};
    const double sg[3] = {sU, sV, sW};
    f[SIFV_F_RHO][dst] = f[SIFV_F_RHO][src];
    for (int d = 0; d < m->dim; ++d) {
        if (varset != SIFV_VARSET_PRIM && f[mom[d]]) f[mom[d]][dst] = sg[d] * f[mom[d]][src];
        if (varset != SIFV_VARSET_CONS && f[vel[d]]) f[vel[d]][dst] = sg[d] * f[vel[d]][src];
    }
    if (varset != SIFV_VARSET_PRIM && f[SIFV_F_RHOE]) f[SIFV_F_RHOE][dst] = f[SIFV_F_RHOE][src];
    if (varset != SIFV_VARSET_CONS && f[SIFV_F_PRES]) f[SIFV_F_PRES][dst] = f[SIFV_F_PRES][src];
    if (f[SIFV_F_SIGMA] && !(exch && varset == SIFV_VARSET_CONS)) f[SIFV_F_SIGMA][dst] = f[SIFV_F_SIGMA][src];
    for (int ph = 0; ph < MAXPH; ++ph) {
        if (f[SIFV_F_ALPHARHO + ph]) f[SIFV_F_ALPHARHO + ph][dst] = f[SIFV_F_ALPHARHO + ph][src];
        if (f[SIFV_F_ALPHA + ph]) f[SIFV_F_ALPHA + ph][dst] = f[SIFV_F_ALPHA + ph][src];
    }
}

/* Source index along one axis for ghost g (1..ng) of a face, and the
 * velocity sign pattern: fillGhostX/Y/Z, src/RectilinearMesh.cpp:137-418;
 * periodic wrap = HaloExchange self exchange, src/HaloExchange.cpp:414-465. */
static int ghost_src(const or_mesh* m, int face, int n, int g, double* sn, double* st) {
    int high = face & 1;
    *sn = 1.0;
    *st = 1.0;
    if (m->wrap[face]) return high ? g - 1 : n - g;
    switch (m->bc[face]) {
        case SIFV_BC_SYMMETRY: return high ? n - g : g - 1;
        case SIFV_BC_SLIPWALL: *sn = -1.0; return high ? n - g : g - 1;
        case SIFV_BC_NOSLIPWALL: *sn = -1.0; *st = -1.0; return high ? n - g : g - 1;
        case SIFV_BC_PERIODIC: return high ? g - 1 : n - g;
        case SIFV_BC_OUTFLOW:
        default: return high ? n - 1 : 0;
    }
}

/* applyBoundaryConditions(state, varSet, halo), src/RectilinearMesh.cpp:498-524 */
static void fill_ghosts(const or_mesh* m, double* const* f, int varset) {
    double sn, st;
    /* x: j,k physical */
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int side = 0; side < 2; ++side)
                for (int g = 1; g <= m->ngx; ++g) {
                    int face = SIFV_XLOW + side;
                    int si = ghost_src(m, face, m->nx, g, &sn, &st);
                    int gi = side ? m->nx - 1 + g : -g;
                    copy_cell(m, f, varset, IDX(m, gi, j, k), IDX(m, si, j, k), sn, st, st, m->mpiwrap[face]);
                }
    if (m->dim >= 2)
        for (int k = 0; k < m->nz; ++k)
            for (int i = -m->ngx; i < m->nx + m->ngx; ++i)
                for (int side = 0; side < 2; ++side)
                    for (int g = 1; g <= m->ngy; ++g) {
                        int face = SIFV_YLOW + side;
                        int sj = ghost_src(m, face, m->ny, g, &sn, &st);
                        int gj = side ? m->ny - 1 + g : -g;
                        copy_cell(m, f, varset, IDX(m, i, gj, k), IDX(m, i, sj, k), st, sn, st, m->mpiwrap[face]);
                    }
    if (m->dim >= 3)
        for (int j = -m->ngy; j < m->ny + m->ngy; ++j)
            for (int i = -m->ngx; i < m->nx + m->ngx; ++i)
                for (int side = 0; side < 2; ++side)
                    for (int g = 1; g <= m->ngz; ++g) {
                        int face = SIFV_ZLOW + side;
                        int sk = ghost_src(m, face, m->nz, g, &sn, &st);
                        int gk = side ? m->nz - 1 + g : -g;
                        copy_cell(m, f, varset, IDX(m, i, j, gk), IDX(m, i, j, sk), st, st, sn, m->mpiwrap[face]);
                    }
}

/* fillScalarGhosts(field, halo), src/RectilinearMesh.cpp:526-624: every
 * non-periodic face is zero-gradient. */
static void fill_scalar(const or_mesh* m, double* a) {
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int g = 1; g <= m->ngx; ++g) {
                a[IDX(m, -g, j, k)] = a[IDX(m, m->wrap[0] ? m->nx - g : 0, j, k)];
                a[IDX(m, m->nx - 1 + g, j, k)] = a[IDX(m, m->wrap[1] ? g - 1 : m->nx - 1, j, k)];
            }
    if (m->dim >= 2)
        for (int k = 0; k < m->nz; ++k)
            for (int i = -m->ngx; i < m->nx + m->ngx; ++i)
                for (int g = 1; g <= m->ngy; ++g) {
                    a[IDX(m, i, -g, k)] = a[IDX(m, i, m->wrap[2] ? m->ny - g : 0, k)];
                    a[IDX(m, i, m->ny - 1 + g, k)] = a[IDX(m, i, m->wrap[3] ? g - 1 : m->ny - 1, k)];
                }
    if (m->dim >= 3)
        for (int j = -m->ngy; j < m->ny + m->ngy; ++j)
            for (int i = -m->ngx; i < m->nx + m->ngx; ++i)
                for (int g = 1; g <= m->ngz; ++g) {
                    a[IDX(m, i, j, -g)] = a[IDX(m, i, j, m->wrap[4] ? m->nz - g : 0)];
                    a[IDX(m, i, j, m->nz - 1 + g)] = a[IDX(m, i, j, m->wrap[5] ? g - 1 : m->nz - 1)];
                }
}

void or_fill_ghosts(or_solver* s, double* const* f, int varset) { fill_ghosts(&s->m, f, varset); }
void or_fill_scalar(or_solver* s, double* a) { fill_scalar(&s->m, a); }

/* ------------------------------------------------------------------ */
/* EOS                                                                  */
/* ------------------------------------------------------------------ */
static int n_phases(const sifv_config* c) { return c->multiPhaseParams.nPhases; }

/* MixtureEOS::effectiveGammaAndPiInf, src/MixtureEOS.cpp:18-31 */
static void eff_gamma_pinf(const double* a, int n, const sifv_config* c, double* g, double* pi) {
    double sInv = 0.0, sPi = 0.0;
    for (int ph = 0; ph < n; ++ph) {
        double gk = c->multiPhaseParams.phases[ph].gamma;
        double gm1 = gk - 1.0;
        sInv += a[ph] / gm1;
        sPi += a[ph] * gk * c->multiPhaseParams.phases[ph].pInf / gm1;
    }
    *g = 1.0 + 1.0 / sInv;
    *pi = (*g - 1.0) / *g * sPi;
}

/* MixtureEOS::mixturePressure, src/MixtureEOS.cpp:35-46 */
static double mix_pressure(double rhoEint, const double* a, int n, const sifv_config* c) {
    double sInv = 0.0, sPi = 0.0;
    for (int ph = 0; ph < n; ++ph) {
        double gm1 = c->multiPhaseParams.phases[ph].gamma - 1.0;
        sInv += a[ph] / gm1;
        sPi += a[ph] * c->multiPhaseParams.phases[ph].gamma * c->multiPhaseParams.phases[ph].pInf / gm1;
    }
    return (rhoEint - sPi) / sInv;
}

/* MixtureEOS::mixtureSoundSpeed, src/MixtureEOS.cpp:48-63 (Wood) */
static double mix_sound(double rho, double p, const double* a, const double* ar, int n,
                        const sifv_config* c) {
    double s = 0.0;
    for (int ph = 0; ph < n; ++ph) {
        double rk = dmax(ar[ph], 1e-14) / dmax(a[ph], 1e-14);
        double gk = c->multiPhaseParams.phases[ph].gamma;
        double pk = c->multiPhaseParams.phases[ph].pInf;
        double ck2 = gk * (p + pk) / rk;
        s += a[ph] / (rk * dmax(ck2, 1e-14));
    }
    double c2 = 1.0 / (rho * dmax(s, 1e-30));
    return sqrt(dmax(c2, 0.0));
}

/* MixtureEOS::mixtureTotalEnergy, src/MixtureEOS.cpp:65-74 */
static double mix_energy(double p, const double* a, int n, double ke, const sifv_config* c) {
    double r = ke;
    for (int ph = 0; ph < n; ++ph) {
        double gm1 = c->multiPhaseParams.phases[ph].gamma - 1.0;
        r += a[ph] * (p + c->multiPhaseParams.phases[ph].gamma * c->multiPhaseParams.phases[ph].pInf) / gm1;
    }
    return r;
}

/* cons->prim on interior cells.
 * Single phase: SolutionState::convertConservativeToPrimitiveVariables
 * (src/SolutionState.cpp:187-205) with IdealGasEOS::toPrimitive
 * (src/IdealGasEOS.cpp:46-64) or StiffenedGasEOS::toPrimitive
 * (src/StiffenedGasEOS.cpp:47-65).
 * Multiphase: MixtureEOS::convertConservativeToPrimitive (src/MixtureEOS.cpp:99-143). */
static void cons_to_prim(or_solver* s, double* const* f, const sifv_config* c) {
    const or_mesh* m = &s->m;
    const int mom[3] = {SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW};
    const int vel[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    int nph = n_phases(c);
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                if (nph > 0) {
                    double r = 0.0;
                    for (int ph = 0; ph < nph; ++ph) r += f[SIFV_F_ALPHARHO + ph][id];
                    f[SIFV_F_RHO][id] = dmax(r, 1e-14);
                    double rs = f[SIFV_F_RHO][id];
                    for (int d = 0; d < m->dim; ++d) f[vel[d]][id] = f[mom[d]][id] / rs;
                    double ke = 0.5 * rs * f[vel[0]][id] * f[vel[0]][id];
                    for (int d = 1; d < m->dim; ++d) ke += 0.5 * rs * f[vel[d]][id] * f[vel[d]][id];
                    double a[MAXPH];
                    for (int ph = 0; ph < nph; ++ph) a[ph] = f[SIFV_F_ALPHA + ph][id];
                    f[SIFV_F_PRES][id] = mix_pressure(f[SIFV_F_RHOE][id] - ke, a, nph, c);
                } else {
                    double wr = dmax(f[SIFV_F_RHO][id], 1e-14);
                    double u[3];
                    for (int d = 0; d < m->dim; ++d) u[d] = f[mom[d]][id] / wr;
                    double ke = 0.0;
                    for (int d = 0; d < m->dim; ++d) ke += u[d] * u[d];
                    ke *= 0.5 * wr;
                    double e = f[SIFV_F_RHOE][id] - ke;
                    double p;
                    if (s->eosKind == SIFV_EOS_STIFFENED)
                        p = (s->gamma - 1.0) * e - s->gamma * s->pInf;
                    else
                        p = dmax((s->gamma - 1.0) * e, 1e-14);
                    for (int d = 0; d < m->dim; ++d) f[vel[d]][id] = u[d];
                    f[SIFV_F_PRES][id] = p;
                }
            }
}

void or_cons_to_prim(or_solver* s, double* const* f, const sifv_config* c) {
    cons_to_prim(s, f, c);
    fill_ghosts(&s->m, f, SIFV_VARSET_PRIM);
}

/* ------------------------------------------------------------------ */
/* reconstruction: src/Reconstruction.cpp:83-200                        */
/* ------------------------------------------------------------------ */
static double weno3L(const double* v, double eps) { /* :83-99 */
    double p0 = -0.5 * v[0] + 1.5 * v[1];
    double p1 = 0.5 * v[1] + 0.5 * v[2];
    double b0 = (v[1] - v[0]) * (v[1] - v[0]);
    double b1 = (v[2] - v[1]) * (v[2] - v[1]);
    double a0 = (1.0 / 3.0) / ((eps + b0) * (eps + b0));
    double a1 = (2.0 / 3.0) / ((eps + b1) * (eps + b1));
    return (a0 * p0 + a1 * p1) / (a0 + a1);
}
static double weno3R(const double* v, double eps) { /* :101-117 */
    double p0 = 0.5 * v[0] + 0.5 * v[1];
    double p1 = 1.5 * v[1] - 0.5 * v[2];
    double b0 = (v[1] - v[0]) * (v[1] - v[0]);
    double b1 = (v[2] - v[1]) * (v[2] - v[1]);
    double a0 = (2.0 / 3.0) / ((eps + b0) * (eps + b0));
    double a1 = (1.0 / 3.0) / ((eps + b1) * (eps + b1));
    return (a0 * p0 + a1 * p1) / (a0 + a1);
}
/* Jiang-Shu WENO5, left-biased: src/Reconstruction.cpp:119-142 */
static double weno5L(const double* v, double eps) {
    double p0 = (1.0 / 3.0) * v[0] - (7.0 / 6.0) * v[1] + (11.0 / 6.0) * v[2];
    double p1 = -(1.0 / 6.0) * v[1] + (5.0 / 6.0) * v[2] + (1.0 / 3.0) * v[3];
    double p2 = (1.0 / 3.0) * v[2] + (5.0 / 6.0) * v[3] - (1.0 / 6.0) * v[4];
    double t0 = v[0] - 2.0 * v[1] + v[2], u0 = v[0] - 4.0 * v[1] + 3.0 * v[2];
    double t1 = v[1] - 2.0 * v[2] + v[3], u1 = v[1] - v[3];
    double t2 = v[2] - 2.0 * v[3] + v[4], u2 = 3.0 * v[2] - 4.0 * v[3] + v[4];
    double b0 = (13.0 / 12.0) * t0 * t0 + (1.0 / 4.0) * u0 * u0;
    double b1 = (13.0 / 12.0) * t1 * t1 + (1.0 / 4.0) * u1 * u1;
    double b2 = (13.0 / 12.0) * t2 * t2 + (1.0 / 4.0) * u2 * u2;
    double a0 = (1.0 / 10.0) / ((eps + b0) * (eps + b0));
    double a1 = (6.0 / 10.0) / ((eps + b1) * (eps + b1));
    double a2 = (3.0 / 10.0) / ((eps + b2) * (eps + b2));
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}
/* mirrored stencil: src/Reconstruction.cpp:144-168 */
static double weno5R(const double* v, double eps) {
    double p0 = (1.0 / 3.0) * v[4] - (7.0 / 6.0) * v[3] + (11.0 / 6.0) * v[2];
    double p1 = -(1.0 / 6.0) * v[3] + (5.0 / 6.0) * v[2] + (1.0 / 3.0) * v[1];
    double p2 = (1.0 / 3.0) * v[2] + (5.0 / 6.0) * v[1] - (1.0 / 6.0) * v[0];
    double t0 = v[4] - 2.0 * v[3] + v[2], u0 = v[4] - 4.0 * v[3] + 3.0 * v[2];
    double t1 = v[3] - 2.0 * v[2] + v[1], u1 = v[3] - v[1];
    double t2 = v[2] - 2.0 * v[1] + v[0], u2 = 3.0 * v[2] - 4.0 * v[1] + v[0];
    double b0 = (13.0 / 12.0) * t0 * t0 + (1.0 / 4.0) * u0 * u0;
    double b1 = (13.0 / 12.0) * t1 * t1 + (1.0 / 4.0) * u1 * u1;
    double b2 = (13.0 / 12.0) * t2 * t2 + (1.0 / 4.0) * u2 * u2;
    double a0 = (1.0 / 10.0) / ((eps + b0) * (eps + b0));
    double a1 = (6.0 / 10.0) / ((eps + b1) * (eps + b1));
    double a2 = (3.0 / 10.0) / ((eps + b2) * (eps + b2));
    return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}
static double upwind3L(const double* v, double e) { /* :170-176 */
    (void)e;
    double p0 = -0.5 * v[0] + 1.5 * v[1];
    double p1 = 0.5 * v[1] + 0.5 * v[2];
    return (1.0 / 3.0) * p0 + (2.0 / 3.0) * p1;
}
static double upwind3R(const double* v, double e) { /* :178-183 */
    (void)e;
    double p0 = 0.5 * v[0] + 0.5 * v[1];
    double p1 = 1.5 * v[1] - 0.5 * v[2];
    return (2.0 / 3.0) * p0 + (1.0 / 3.0) * p1;
}
static double upwind5L(const double* v, double e) { /* :185-192 */
    (void)e;
    double p0 = (1.0 / 3.0) * v[0] - (7.0 / 6.0) * v[1] + (11.0 / 6.0) * v[2];
    double p1 = -(1.0 / 6.0) * v[1] + (5.0 / 6.0) * v[2] + (1.0 / 3.0) * v[3];
    double p2 = (1.0 / 3.0) * v[2] + (5.0 / 6.0) * v[3] - (1.0 / 6.0) * v[4];
    return (1.0 / 10.0) * p0 + (6.0 / 10.0) * p1 + (3.0 / 10.0) * p2;
}
static double upwind5R(const double* v, double e) { /* :194-200 */
    (void)e;
    double p0 = (1.0 / 3.0) * v[4] - (7.0 / 6.0) * v[3] + (11.0 / 6.0) * v[2];
    double p1 = -(1.0 / 6.0) * v[3] + (5.0 / 6.0) * v[2] + (1.0 / 3.0) * v[1];
    double p2 = (1.0 / 3.0) * v[2] + (5.0 / 6.0) * v[1] - (1.0 / 6.0) * v[0];
    return (1.0 / 10.0) * p0 + (6.0 / 10.0) * p1 + (3.0 / 10.0) * p2;
}

double or_recon(int order, int side, const double* v, double eps) {
    switch (order) {
        case SIFV_WENO3: return side ? weno3R(v, eps) : weno3L(v, eps);
        case SIFV_WENO5: return side ? weno5R(v, eps) : weno5L(v, eps);
        case SIFV_UPWIND3: return side ? upwind3R(v, eps) : upwind3L(v, eps);
        case SIFV_UPWIND5: return side ? upwind5R(v, eps) : upwind5L(v, eps);
        default: return v[0];
    }
}

/* Face states for one direction: Reconstructor::reconstructX/Y/Z
 * (src/Reconstruction.cpp:204-541).  dir = 0,1,2; face f between cell
 * (c - e_dir) and c.  Fields read: rho, vel*, pres, sigma, alpha. */
static void recon_face(const or_solver* s, double* const* f, const sifv_config* c, int dir,
                       int i, int j, int k, prim_t* L, prim_t* R) {
    const or_mesh* m = &s->m;
    int order = c->reconOrder;
    int nph = n_phases(c);
    memset(L, 0, sizeof *L);
    memset(R, 0, sizeof *R);
    int di = dir == 0, dj = dir == 1, dk = dir == 2;
    int r = (order == SIFV_WENO5 || order == SIFV_UPWIND5) ? 3
          : (order == SIFV_WENO3 || order == SIFV_UPWIND3) ? 2 : 1;
    size_t cell[6];
    for (int q = 0; q < 2 * r; ++q) {
        int o = q - r; /* offsets -r .. r-1 */
        cell[q] = IDX(m, i + o * di, j + o * dj, k + o * dk);
    }
    const double* var[5 + 2 * MAXPH];
    double* outL[5 + 2 * MAXPH];
    double* outR[5 + 2 * MAXPH];
    int nv = 0;
    var[nv] = f[SIFV_F_RHO]; outL[nv] = &L->rho; outR[nv] = &R->rho; ++nv;
    const int vel[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    for (int d = 0; d < m->dim; ++d) {
        var[nv] = f[vel[d]]; outL[nv] = &L->u[d]; outR[nv] = &R->u[d]; ++nv;
    }
    var[nv] = f[SIFV_F_PRES]; outL[nv] = &L->p; outR[nv] = &R->p; ++nv;
    var[nv] = f[SIFV_F_SIGMA]; outL[nv] = &L->sigma; outR[nv] = &R->sigma; ++nv;
    for (int ph = 0; ph < nph; ++ph) {
        var[nv] = f[SIFV_F_ALPHA + ph]; outL[nv] = &L->alpha[ph]; outR[nv] = &R->alpha[ph]; ++nv;
    }
    for (int q = 0; q < nv; ++q) {
        const double* a = var[q];
        if (r == 1) {
            *outL[q] = a[cell[0]];
            *outR[q] = a[cell[1]];
        } else {
            double vl[5], vr[5];
            int sz = 2 * r - 1;
            for (int t = 0; t < sz; ++t) {
                vl[t] = a[cell[t]];
                vr[t] = a[cell[t + 1]];
            }
            *outL[q] = or_recon(order, 0, vl, c->wenoEps);
            *outR[q] = or_recon(order, 1, vr, c->wenoEps);
        }
    }
    if (nph > 0) {
        eff_gamma_pinf(L->alpha, nph, c, &L->gammaEff, &L->piInfEff);
        eff_gamma_pinf(R->alpha, nph, c, &R->gammaEff, &R->piInfEff);
    } else {
        L->gammaEff = s->gamma;
        L->piInfEff = s->pInf;
        R->gammaEff = s->gamma;
        R->piInfEff = s->pInf;
    }
}

/* ------------------------------------------------------------------ */
/* Riemann fluxes                                                       */
/* ------------------------------------------------------------------ */
/* soundSpeedDirect / rhoEFromState, include/RiemannSolver.hpp:69-78 */
static double sound_direct(const prim_t* w) {
    return sqrt(w->gammaEff * dmax(w->p + w->piInfEff, 1e-14) / dmax(w->rho, 1e-14));
}
static double rhoE_state(const prim_t* w) {
    double ke = 0.5 * w->rho * (w->u[0] * w->u[0] + w->u[1] * w->u[1] + w->u[2] * w->u[2]);
    return (w->p + w->gammaEff * w->piInfEff) / (w->gammaEff - 1.0) + ke;
}
static double vnormal(const prim_t* w, const double* n, int dim) { /* :58-63 */
    double vn = 0.0;
    for (int d = 0; d < dim; ++d) vn += w->u[d] * n[d];
    return vn;
}

/* computeHLLCFlux, src/HLLCSolver.cpp:7-161 */
static void hllc(const prim_t* L, const prim_t* R, const double* n, int dim, int incP, int nph,
                 flux_t* F) {
    memset(F, 0, sizeof *F);
    double uL = vnormal(L, n, dim), uR = vnormal(R, n, dim);
    double EL = rhoE_state(L), ER = rhoE_state(R);
    double sL, sR, sS;
    if (incP) {
        double cL = sound_direct(L), cR = sound_direct(R);
        sL = dmin(uL - cL, uR - cR);
        sR = dmax(uL + cL, uR + cR);
        sS = (R->p - L->p + L->rho * uL * (sL - uL) - R->rho * uR * (sR - uR)) /
             (L->rho * (sL - uL) - R->rho * (sR - uR));
    } else {
        sL = dmin(uL, uR);
        sR = dmax(uL, uR);
        sS = 0.5 * (uL + uR);
    }
    if (sL >= 0) {
        F->mass = L->rho * uL;
        for (int i = 0; i < dim; ++i) F->mom[i] = L->rho * L->u[i] * uL;
        F->energy = EL * uL;
        if (incP) {
            for (int i = 0; i < dim; ++i) F->mom[i] += L->p * n[i];
            F->energy += L->p * uL;
        }
    } else if (sR <= 0) {
        F->mass = R->rho * uR;
        for (int i = 0; i < dim; ++i) F->mom[i] = R->rho * R->u[i] * uR;
        F->energy = ER * uR;
        if (incP) {
            for (int i = 0; i < dim; ++i) F->mom[i] += R->p * n[i];
            F->energy += R->p * uR;
        }
    } else if (sS >= 0) {
        double rS = L->rho * (sL - uL) / (sL - sS);
        F->mass = L->rho * uL + sL * (rS - L->rho);
        if (incP) {
            for (int i = 0; i < dim; ++i) {
                double ruS = rS * (L->u[i] + (sS - uL) * n[i]);
                F->mom[i] = L->rho * L->u[i] * uL + L->p * n[i] + sL * (ruS - L->rho * L->u[i]);
            }
            double eL = EL / L->rho;
            double ES = rS * (eL + (sS - uL) * (sS + L->p / (L->rho * (sL - uL))));
            F->energy = (EL + L->p) * uL + sL * (ES - EL);
        } else {
            for (int i = 0; i < dim; ++i) {
                double ruS = rS * (L->u[i] + (sS - uL) * n[i]);
                F->mom[i] = L->rho * L->u[i] * uL + sL * (ruS - L->rho * L->u[i]);
            }
            double eL = EL / L->rho;
            double ES = rS * (eL + (sS - uL) * sS);
            F->energy = EL * uL + sL * (ES - EL);
        }
    } else {
        double rS = R->rho * (sR - uR) / (sR - sS);
        F->mass = R->rho * uR + sR * (rS - R->rho);
        if (incP) {
            for (int i = 0; i < dim; ++i) {
                double ruS = rS * (R->u[i] + (sS - uR) * n[i]);
                F->mom[i] = R->rho * R->u[i] * uR + R->p * n[i] + sR * (ruS - R->rho * R->u[i]);
            }
            double eR = ER / R->rho;
            double ES = rS * (eR + (sS - uR) * (sS + R->p / (R->rho * (sR - uR))));
            F->energy = (ER + R->p) * uR + sR * (ES - ER);
        } else {
            for (int i = 0; i < dim; ++i) {
                double ruS = rS * (R->u[i] + (sS - uR) * n[i]);
                F->mom[i] = R->rho * R->u[i] * uR + sR * (ruS - R->rho * R->u[i]);
            }
            double eR = ER / R->rho;
            double ES = rS * (eR + (sS - uR) * sS);
            F->energy = ER * uR + sR * (ES - ER);
        }
    }
    /* contact-type quantities: src/HLLCSolver.cpp:137-158 */
    if (sL >= 0) {
        F->faceVelocity = uL;
        F->pressureFlux = L->p * uL;
        for (int ph = 0; ph < nph; ++ph) F->alphaFlux[ph] = L->alpha[ph] * uL;
    } else if (sR <= 0) {
        F->faceVelocity = uR;
        F->pressureFlux = R->p * uR;
        for (int ph = 0; ph < nph; ++ph) F->alphaFlux[ph] = R->alpha[ph] * uR;
    } else {
        F->faceVelocity = sS;
        const prim_t* up = (sS >= 0) ? L : R;
        F->pressureFlux = up->p * sS;
        for (int ph = 0; ph < nph; ++ph) F->alphaFlux[ph] = up->alpha[ph] * sS;
    }
}

/* computeRusanovFlux, src/RusanovSolver.cpp:7-83 */
static void rusanov(const prim_t* L, const prim_t* R, const double* n, int dim, int incP, int nph,
                    flux_t* F) {
    memset(F, 0, sizeof *F);
    double uL = vnormal(L, n, dim), uR = vnormal(R, n, dim);
    double aL = fabs(uL), aR = fabs(uR);
    double sM;
    if (incP) {
        double cL = sound_direct(L), cR = sound_direct(R);
        sM = dmax(aL + cL, aR + cR);
    } else {
        sM = dmax(aL, aR);
    }
    double EL = rhoE_state(L), ER = rhoE_state(R);
    double mL = L->rho * uL, mR = R->rho * uR;
    double pL[3] = {0, 0, 0}, pR[3] = {0, 0, 0};
    for (int i = 0; i < dim; ++i) {
        pL[i] = L->rho * L->u[i] * uL;
        pR[i] = R->rho * R->u[i] * uR;
    }
    double eFL = EL * uL, eFR = ER * uR;
    if (incP) {
        for (int i = 0; i < dim; ++i) {
            pL[i] += L->p * n[i];
            pR[i] += R->p * n[i];
        }
        eFL += L->p * uL;
        eFR += R->p * uR;
    }
    F->mass = 0.5 * (mL + mR) - 0.5 * sM * (R->rho - L->rho);
    for (int i = 0; i < dim; ++i)
        F->mom[i] = 0.5 * (pL[i] + pR[i]) - 0.5 * sM * (R->rho * R->u[i] - L->rho * L->u[i]);
    F->energy = 0.5 * (eFL + eFR) - 0.5 * sM * (ER - EL);
    F->faceVelocity = 0.5 * (uL + uR);
    F->pressureFlux = 0.5 * (L->p * uL + R->p * uR) - 0.5 * sM * (R->p - L->p);
    for (int ph = 0; ph < nph; ++ph)
        F->alphaFlux[ph] = 0.5 * (L->alpha[ph] * uL + R->alpha[ph] * uR) -
                           0.5 * sM * (R->alpha[ph] - L->alpha[ph]);
}

/* computeLFFlux, src/LFSolver.cpp:8-75 */
static void lf(const prim_t* L, const prim_t* R, const double* n, int dim, int incP, int igr,
               int nph, flux_t* F) {
    memset(F, 0, sizeof *F);
    double uL = vnormal(L, n, dim), uR = vnormal(R, n, dim);
    double qL = 0.0, qR = 0.0;
    for (int i = 0; i < dim; ++i) {
        qL += L->u[i] * L->u[i];
        qR += R->u[i] * R->u[i];
    }
    qL = sqrt(qL);
    qR = sqrt(qR);
    double C;
    if (incP) {
        double cL = sound_direct(L), cR = sound_direct(R);
        C = dmax(qL, qR) + dmax(cL, cR);
    } else {
        C = dmax(qL, qR);
    }
    F->mass = 0.5 * (L->rho * uL + R->rho * uR) - 0.5 * C * (R->rho - L->rho);
    for (int i = 0; i < dim; ++i)
        F->mom[i] = 0.5 * (L->rho * L->u[i] * uL + R->rho * R->u[i] * uR) -
                    0.5 * C * (R->rho * R->u[i] - L->rho * L->u[i]);
    double EL = rhoE_state(L), ER = rhoE_state(R);
    F->energy = 0.5 * (EL * uL + ER * uR) - 0.5 * C * (ER - EL);
    if (incP) {
        for (int i = 0; i < dim; ++i) F->mom[i] += 0.5 * (L->p + R->p) * n[i];
        F->energy += 0.5 * (L->p * uL + R->p * uR);
    }
    if (igr) {
        for (int i = 0; i < dim; ++i) F->mom[i] += 0.5 * (L->sigma + R->sigma) * n[i];
        F->energy += 0.5 * (L->sigma * uL + R->sigma * uR);
    }
    F->faceVelocity = 0.5 * (uL + uR);
    F->pressureFlux = 0.5 * (L->p * uL + R->p * uR) - 0.5 * C * (R->p - L->p);
    for (int ph = 0; ph < nph; ++ph)
        F->alphaFlux[ph] = 0.5 * (L->alpha[ph] * uL + R->alpha[ph] * uR) -
                           0.5 * C * (R->alpha[ph] - L->alpha[ph]);
}

/* computeFluxDirect, include/RiemannSolver.hpp:91-100 */
static void flux_direct(int type, const prim_t* L, const prim_t* R, const double* n, int dim,
                        int incP, int igr, int nph, flux_t* F) {
    switch (type) {
        case SIFV_RUSANOV: rusanov(L, R, n, dim, incP, nph, F); break;
        case SIFV_HLLC: hllc(L, R, n, dim, incP, nph, F); break;
        default: lf(L, R, n, dim, incP, igr, nph, F); break;
    }
}

static void unpack_prim(const double* a, prim_t* w) {
    w->rho = a[0];
    w->u[0] = a[1];
    w->u[1] = a[2];
    w->u[2] = a[3];
    w->p = a[4];
    w->sigma = a[5];
    w->gammaEff = a[6];
    w->piInfEff = a[7];
    for (int ph = 0; ph < MAXPH; ++ph) w->alpha[ph] = a[8 + ph];
}

void or_riemann_flux(int type, const double* Lp, const double* Rp, int dir, int dim, int incP,
                     int igr, int nph, double* out) {
    prim_t L, R;
    flux_t F;
    unpack_prim(Lp, &L);
    unpack_prim(Rp, &R);
    double n[3] = {0, 0, 0};
    n[dir] = 1.0;
    flux_direct(type, &L, &R, n, dim, incP, igr, nph, &F);
    out[0] = F.mass;
    out[1] = F.mom[0];
    out[2] = F.mom[1];
    out[3] = F.mom[2];
    out[4] = F.energy;
    for (int ph = 0; ph < MAXPH; ++ph) out[5 + ph] = F.alphaFlux[ph];
    out[13] = F.pressureFlux;
    out[14] = F.faceVelocity;
}

/* ------------------------------------------------------------------ */
/* RHS: ExplicitSolver::computeRHS (src/ExplicitSolver.cpp:270-496) and   */
/* SemiImplicitSolver::computeRHS (src/SemiImplicitSolver.cpp:238-482)    */
/* ------------------------------------------------------------------ */
/* viscous stress: addViscousFluxes (src/ViscousFlux.cpp:25-303).  For the
 * face between cells L and R along `dir`: normal derivatives from L and R,
 * transverse derivatives as the mean of the central differences in L and R,
 * tau = mu (grad u + grad u^T) - 2/3 mu div(u) I on the face row, work
 * tau . u_face; L gains +area/vol(L) * (tau, work), R gains -area/vol(R).
 * Faces are visited in the reference's loop order (x, y, z; k, j, i).  Face
 * viscosity: the scalar mu, or the mean of sum_k alpha_k mu_k of L and R. */
static double visc_face_mu(double* const* f, const sifv_config* c, size_t L, size_t R) {
    if (c->viscousNPhaseMu <= 0) return c->viscousMu;
    double mL = 0.0, mR = 0.0;
    for (int ph = 0; ph < c->viscousNPhaseMu && ph < c->multiPhaseParams.nPhases; ++ph) {
        mL += f[SIFV_F_ALPHA + ph][L] * c->viscousPhaseMu[ph];
        mR += f[SIFV_F_ALPHA + ph][R] * c->viscousPhaseMu[ph];
    }
    return 0.5 * (mL + mR);
}

static void add_viscous(or_solver* s, double* const* f, const sifv_config* c) {
    or_mesh* m = &s->m;
    const int dim = m->dim;
    const int velId[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    const int momId[3] = {SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW};
    for (int dir = 0; dir < dim; ++dir) {
        const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
        for (int k = 0; k < m->nz + ek; ++k)
            for (int j = 0; j < m->ny + ej; ++j)
                for (int i = 0; i < m->nx + ei; ++i) {
                    const int li = i - ei, lj = j - ej, lk = k - ek;
                    const size_t L = IDX(m, li, lj, lk), R = IDX(m, i, j, k);
                    /* grad[a][b] = d u_a / d x_b on the face */
                    double grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                    double hn = dir == 0 ? 0.5 * (DX(m, li) + DX(m, i))
                                         : dir == 1 ? 0.5 * (DY(m, lj) + DY(m, j)) : 0.5 * (DZ(m, lk) + DZ(m, k));
                    for (int a = 0; a < dim; ++a) grad[a][dir] = (f[velId[a]][R] - f[velId[a]][L]) / hn;
                    for (int b = 0; b < dim; ++b) {
                        if (b == dir) continue;
                        const int bi = b == 0, bj = b == 1, bk = b == 2;
                        /* transverse spacing of the face's row: the R cell's width along b */
                        double hb = b == 0 ? DX(m, i) : b == 1 ? DY(m, j) : DZ(m, k);
                        const size_t Lm = IDX(m, li - bi, lj - bj, lk - bk), Lp = IDX(m, li + bi, lj + bj, lk + bk);
                        const size_t Rm = IDX(m, i - bi, j - bj, k - bk), Rp = IDX(m, i + bi, j + bj, k + bk);
                        for (int a = 0; a < dim; ++a) {
                            /* the reference forms only the components its stress needs */
                            if (!(a == b || a == dir)) continue;
                            double dL = (f[velId[a]][Lp] - f[velId[a]][Lm]) / (2.0 * hb);
                            double dR = (f[velId[a]][Rp] - f[velId[a]][Rm]) / (2.0 * hb);
                            grad[a][b] = 0.5 * (dL + dR);
                        }
                    }
                    double divU = grad[0][0];
                    if (dim >= 2) divU += grad[1][1];
                    if (dim >= 3) divU += grad[2][2];
                    const double mu = visc_face_mu(f, c, L, R);
                    double tau[3] = {0, 0, 0};
                    for (int a = 0; a < 3; ++a) {
                        if (a == dir) tau[a] = mu * (2.0 * grad[dir][dir] - (2.0 / 3.0) * divU);
                        else tau[a] = mu * (grad[a][dir] + grad[dir][a]);
                    }
                    double uf[3] = {0, 0, 0};
                    for (int a = 0; a < dim; ++a) uf[a] = 0.5 * (f[velId[a]][L] + f[velId[a]][R]);
                    double work = tau[0] * uf[0] + tau[1] * uf[1] + tau[2] * uf[2];
                    double area = dir == 0 ? DY(m, j) * DZ(m, k) : dir == 1 ? DX(m, i) * DZ(m, k) : DX(m, i) * DY(m, j);
                    const int inL = dir == 0 ? i >= 1 : dir == 1 ? j >= 1 : k >= 1;
                    const int inR = dir == 0 ? i < m->nx : dir == 1 ? j < m->ny : k < m->nz;
                    if (inL) {
                        double coeff = area / VOL(m, li, lj, lk);
                        for (int a = 0; a < dim; ++a) s->rhs[momId[a]][L] += coeff * tau[a];
                        s->rhs[SIFV_F_RHOE][L] += coeff * work;
                    }
                    if (inR) {
                        double coeff = area / VOL(m, i, j, k);
                        for (int a = 0; a < dim; ++a) s->rhs[momId[a]][R] -= coeff * tau[a];
                        s->rhs[SIFV_F_RHOE][R] -= coeff * work;
                    }
                }
    }
}

/* surface tension: addSurfaceTensionFluxes (src/SurfaceTension.cpp:7-255).
 * Capillary stress of the face between L and R along `dir` from the
 * gradient g of alpha[0] on the face (normal component from L and R,
 * transverse components the mean of the central differences in L and R):
 * T[dir] = sigma (|g|^2 - g_dir^2) / max(|g|, eps), T[a] = -sigma g_dir g_a /
 * max(|g|, eps) (a != dir), work T . u_face; L gains +area/vol(L) * (T, work),
 * R gains -area/vol(R).  Faces in the reference's loop order. */
static void add_surface_tension(or_solver* s, double* const* f, const sifv_config* c) {
    or_mesh* m = &s->m;
    const int dim = m->dim;
    const double sigma = c->surfaceTensionSigma, eps = c->surfaceTensionEps;
    const double* al = f[SIFV_F_ALPHA];
    const int velId[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    const int momId[3] = {SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW};
    for (int dir = 0; dir < dim; ++dir) {
        const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
        for (int k = 0; k < m->nz + ek; ++k)
            for (int j = 0; j < m->ny + ej; ++j)
                for (int i = 0; i < m->nx + ei; ++i) {
                    const int li = i - ei, lj = j - ej, lk = k - ek;
                    const size_t L = IDX(m, li, lj, lk), R = IDX(m, i, j, k);
                    double g[3] = {0, 0, 0};
                    double hn = dir == 0 ? 0.5 * (DX(m, li) + DX(m, i))
                                         : dir == 1 ? 0.5 * (DY(m, lj) + DY(m, j)) : 0.5 * (DZ(m, lk) + DZ(m, k));
                    g[dir] = (al[R] - al[L]) / hn;
                    for (int b = 0; b < dim; ++b) {
                        if (b == dir) continue;
                        const int bi = b == 0, bj = b == 1, bk = b == 2;
                        double hb = b == 0 ? DX(m, i) : b == 1 ? DY(m, j) : DZ(m, k);
                        const size_t Lm = IDX(m, li - bi, lj - bj, lk - bk), Lp = IDX(m, li + bi, lj + bj, lk + bk);
                        const size_t Rm = IDX(m, i - bi, j - bj, k - bk), Rp = IDX(m, i + bi, j + bj, k + bk);
                        double dL = (al[Lp] - al[Lm]) / (2.0 * hb);
                        double dR = (al[Rp] - al[Rm]) / (2.0 * hb);
                        g[b] = 0.5 * (dL + dR);
                    }
                    /* |g| summed in x, y, z order */
                    double ng = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                    double nr = dmax(ng, eps);
                    double T[3];
                    for (int a = 0; a < 3; ++a) {
                        if (a == dir) {
                            /* the squares of the two transverse components, in axis order */
                            const int b0 = dir == 0 ? 1 : 0, b1 = dir == 2 ? 1 : 2;
                            T[a] = sigma * (g[b0] * g[b0] + g[b1] * g[b1]) / nr;
                        } else {
                            T[a] = -sigma * g[dir] * g[a] / nr;
                        }
                    }
                    double uf[3] = {0, 0, 0};
                    for (int a = 0; a < dim; ++a) uf[a] = 0.5 * (f[velId[a]][L] + f[velId[a]][R]);
                    double work = T[0] * uf[0] + T[1] * uf[1] + T[2] * uf[2];
                    double area = dir == 0 ? DY(m, j) * DZ(m, k) : dir == 1 ? DX(m, i) * DZ(m, k) : DX(m, i) * DY(m, j);
                    const int inL = dir == 0 ? i >= 1 : dir == 1 ? j >= 1 : k >= 1;
                    const int inR = dir == 0 ? i < m->nx : dir == 1 ? j < m->ny : k < m->nz;
                    if (inL) {
                        double coeff = area / VOL(m, li, lj, lk);
                        for (int a = 0; a < dim; ++a) s->rhs[momId[a]][L] += coeff * T[a];
                        s->rhs[SIFV_F_RHOE][L] += coeff * work;
                    }
                    if (inR) {
                        double coeff = area / VOL(m, i, j, k);
                        for (int a = 0; a < dim; ++a) s->rhs[momId[a]][R] -= coeff * T[a];
                        s->rhs[SIFV_F_RHOE][R] -= coeff * work;
                    }
                }
    }
}

/* computeCapillaryDt (src/RKTimeStepping.cpp:201-228) */
static double capillary_dt(or_solver* s, double* const* f, double sigma, double cfl, double maxDt) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    double dt = maxDt;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                double h = DX(m, i);
                if (dim >= 2) h = dmin(h, DY(m, j));
                if (dim >= 3) h = dmin(h, DZ(m, k));
                double dtCell = sqrt(dmax(f[SIFV_F_RHO][id], 1e-14) * h * h * h / sigma);
                dt = dmin(dt, cfl * dtCell);
            }
    return dt;
}

/* computeViscousDt (src/RKTimeStepping.cpp:161-188): scalar mu only */
static double viscous_dt(or_solver* s, double* const* f, double mu, double cfl, double maxDt) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    double dt = maxDt;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                double h = DX(m, i);
                if (dim >= 2) h = dmin(h, DY(m, j));
                if (dim >= 3) h = dmin(h, DZ(m, k));
                double nu = mu / dmax(f[SIFV_F_RHO][id], 1e-14);
                double dtCell = h * h / (2.0 * dim * nu);
                dt = dmin(dt, cfl * dtCell);
            }
    return dt;
}

static int has_viscosity(const sifv_config* c) { return c->viscousMu > 0.0 || c->viscousNPhaseMu > 0; }

static void compute_rhs(or_solver* s, double* const* f, const sifv_config* c, int semi) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    int nph = n_phases(c);
    int incP = !c->semiImplicit;
    size_t n = m->total;
    const int cons[5] = {SIFV_F_RHO, SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW, SIFV_F_RHOE};
    for (int q = 0; q < 5; ++q) memset(s->rhs[cons[q]], 0, n * sizeof(double));
    for (int ph = 0; ph < nph; ++ph) {
        memset(s->rhs[SIFV_F_ALPHARHO + ph], 0, n * sizeof(double));
        memset(s->rhs[SIFV_F_ALPHA + ph], 0, n * sizeof(double));
    }
    memset(s->rhsP, 0, n * sizeof(double));
    memset(s->divU, 0, n * sizeof(double));
    int trackDiv = semi || nph > 0;
    for (int dir = 0; dir < dim; ++dir) {
        int ex = dir == 0, ey = dir == 1, ez = dir == 2;
        double nrm[3] = {0, 0, 0};
        nrm[dir] = 1.0;
        int nxF = m->nx + ex, nyF = m->ny + ey, nzF = m->nz + ez;
        for (int k = 0; k < nzF; ++k)
            for (int j = 0; j < nyF; ++j)
                for (int i = 0; i < nxF; ++i) {
                    prim_t L, R;
                    flux_t F;
                    recon_face(s, f, c, dir, i, j, k, &L, &R);
                    flux_direct(s->riemann, &L, &R, nrm, dim, incP, c->useIGR, nph, &F);
                    double area = dir == 0 ? DY(m, j) * DZ(m, k)
                                : dir == 1 ? DX(m, i) * DZ(m, k) : DX(m, i) * DY(m, j);
                    int li = i - ex, lj = j - ey, lk = k - ez;
                    size_t up = 0;
                    if (nph > 0) up = (F.mass >= 0) ? IDX(m, li, lj, lk) : IDX(m, i, j, k);
                    int pos = dir == 0 ? i : dir == 1 ? j : k;
                    int npos = dir == 0 ? m->nx : dir == 1 ? m->ny : m->nz;
                    for (int side = 0; side < 2; ++side) {
                        int ci = side ? i : li, cj = side ? j : lj, ck = side ? k : lk;
                        if (side == 0 && pos < 1) continue;
                        if (side == 1 && pos >= npos) continue;
                        size_t id = IDX(m, ci, cj, ck);
                        double coeff = area / VOL(m, ci, cj, ck);
                        double sg = side ? 1.0 : -1.0;
                        /* rhs[L] -= coeff*F ; rhs[R] += coeff*F */
                        if (side) {
                            s->rhs[SIFV_F_RHO][id] += coeff * F.mass;
                            s->rhs[SIFV_F_RHOU][id] += coeff * F.mom[0];
                            if (dim >= 2) s->rhs[SIFV_F_RHOV][id] += coeff * F.mom[1];
                            if (dim >= 3) s->rhs[SIFV_F_RHOW][id] += coeff * F.mom[2];
                            s->rhs[SIFV_F_RHOE][id] += coeff * F.energy;
                        } else {
                            s->rhs[SIFV_F_RHO][id] -= coeff * F.mass;
                            s->rhs[SIFV_F_RHOU][id] -= coeff * F.mom[0];
                            if (dim >= 2) s->rhs[SIFV_F_RHOV][id] -= coeff * F.mom[1];
                            if (dim >= 3) s->rhs[SIFV_F_RHOW][id] -= coeff * F.mom[2];
                            s->rhs[SIFV_F_RHOE][id] -= coeff * F.energy;
                        }
                        if (semi) {
                            if (side) s->rhsP[id] += coeff * F.pressureFlux;
                            else s->rhsP[id] -= coeff * F.pressureFlux;
                        }
                        if (trackDiv) {
                            if (side) s->divU[id] -= coeff * F.faceVelocity;
                            else s->divU[id] += coeff * F.faceVelocity;
                        }
                        if (nph > 0) {
                            double ru = dmax(f[SIFV_F_RHO][up], 1e-14);
                            for (int ph = 0; ph < nph; ++ph) {
                                double arF = (f[SIFV_F_ALPHARHO + ph][up] / ru) * F.mass;
                                if (side) s->rhs[SIFV_F_ALPHARHO + ph][id] += coeff * arF;
                                else s->rhs[SIFV_F_ALPHARHO + ph][id] -= coeff * arF;
                            }
                            for (int ph = 0; ph < nph; ++ph) {
                                if (side) s->rhs[SIFV_F_ALPHA + ph][id] += coeff * F.alphaFlux[ph];
                                else s->rhs[SIFV_F_ALPHA + ph][id] -= coeff * F.alphaFlux[ph];
                            }
                        }
                        (void)sg;
                    }
                }
    }
    /* alpha source, src/ExplicitSolver.cpp:486-496 */
    if (nph > 0)
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    for (int ph = 0; ph < nph; ++ph)
                        s->rhs[SIFV_F_ALPHA + ph][id] += f[SIFV_F_ALPHA + ph][id] * s->divU[id];
                }
    /* p div(u) source, src/SemiImplicitSolver.cpp:472-482 */
    if (semi)
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    s->rhsP[id] += f[SIFV_F_PRES][id] * s->divU[id];
                }
    /* body force: a + b cos(c t + d) per axis at the config time, momentum
     * rho*a and energy rho*(u.a) (src/ExplicitSolver.cpp:498-521,
     * src/SemiImplicitSolver.cpp:484-506; hasBodyForce,
     * include/SimulationConfig.hpp:101-106) */
    int bf = 0;
    for (int q = 0; q < 3; ++q)
        if (c->bodyForceParams.a[q] != 0.0 || c->bodyForceParams.b[q] != 0.0) bf = 1;
    if (bf) {
        double acc[3];
        for (int q = 0; q < 3; ++q)
            acc[q] = c->bodyForceParams.a[q] + c->bodyForceParams.b[q] * cos(c->bodyForceParams.c[q] * c->time +
                                                                              c->bodyForceParams.d[q]);
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double rho = f[SIFV_F_RHO][id];
                    s->rhs[SIFV_F_RHOU][id] += rho * acc[0];
                    if (dim >= 2) s->rhs[SIFV_F_RHOV][id] += rho * acc[1];
                    if (dim >= 3) s->rhs[SIFV_F_RHOW][id] += rho * acc[2];
                    double work = f[SIFV_F_VELU][id] * acc[0];
                    if (dim >= 2) work += f[SIFV_F_VELV][id] * acc[1];
                    if (dim >= 3) work += f[SIFV_F_VELW][id] * acc[2];
                    s->rhs[SIFV_F_RHOE][id] += rho * work;
                }
    }
    if (has_viscosity(c)) add_viscous(s, f, c);
    if (c->surfaceTensionSigma > 0.0) add_surface_tension(s, f, c);
}

/* ------------------------------------------------------------------ */
/* IGR: src/IGR.cpp:18-200, solver gradient loop src/ExplicitSolver.cpp:215-267 */
/* ------------------------------------------------------------------ */
static void velocity_gradients(or_solver* s, double* const* f) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    const double* U = f[SIFV_F_VELU];
    const double* V = dim >= 2 ? f[SIFV_F_VELV] : NULL;
    const double* W = dim >= 3 ? f[SIFV_F_VELW] : NULL;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                double* g = s->grad + 9 * id;
                for (int q = 0; q < 9; ++q) g[q] = 0.0;
                size_t xm = IDX(m, i - 1, j, k), xp = IDX(m, i + 1, j, k);
                double invDx = 0.5 / DX(m, i);
                g[0] = (U[xp] - U[xm]) * invDx; /* g[a*3+b] = du_a/dx_b */
                if (dim >= 2) {
                    size_t ym = IDX(m, i, j - 1, k), yp = IDX(m, i, j + 1, k);
                    double invDy = 0.5 / DY(m, j);
                    g[3] = (V[xp] - V[xm]) * invDx;
                    g[1] = (U[yp] - U[ym]) * invDy;
                    g[4] = (V[yp] - V[ym]) * invDy;
                    if (dim >= 3) {
                        size_t zm = IDX(m, i, j, k - 1), zp = IDX(m, i, j, k + 1);
                        double invDz = 0.5 / DZ(m, k);
                        g[7] = (W[yp] - W[ym]) * invDy;
                        g[6] = (W[xp] - W[xm]) * invDx;
                        g[2] = (U[zp] - U[zm]) * invDz;
                        g[5] = (V[zp] - V[zm]) * invDz;
                        g[8] = (W[zp] - W[zm]) * invDz;
                    }
                }
            }
}

/* IGRSolver::solveEntropicPressure (halo variant), src/IGR.cpp:102-158 */
static void igr_solve(or_solver* s, double* const* f, const sifv_config* c) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    velocity_gradients(s, f);
    int iters = c->step == 0 ? c->igrParams.IGRWarmStartIters : c->igrParams.IGRIters;
    double alpha = c->igrParams.alphaCoeff * DX(m, 0) * DX(m, 0);
    double* rho = f[SIFV_F_RHO];
    double* sg = f[SIFV_F_SIGMA];
    for (int it = 0; it < iters; ++it) {
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    const double* g = s->grad + 9 * id;
                    /* computeIGRRhs, src/IGR.cpp:18-35 */
                    double trSq = 0.0;
                    for (int a = 0; a < dim; ++a)
                        for (int b = 0; b < dim; ++b) trSq += g[3 * a + b] * g[3 * b + a];
                    double tr = 0.0;
                    for (int a = 0; a < dim; ++a) tr += g[3 * a + a];
                    tr *= tr;
                    double rhs = alpha * (trSq + tr);
                    double diag = 1.0 / rho[id];
                    double off = 0.0;
                    double ri = rho[id];
                    {
                        size_t l = IDX(m, i - 1, j, k), r = IDX(m, i + 1, j, k);
                        double fl = 0.5 * (ri + rho[l]), fr = 0.5 * (ri + rho[r]);
                        double d2 = 1.0 / (DX(m, i) * DX(m, i));
                        diag += alpha * d2 * (1.0 / fl + 1.0 / fr);
                        off += alpha * d2 * (sg[l] / fl + sg[r] / fr);
                    }
                    if (dim >= 2) {
                        size_t l = IDX(m, i, j - 1, k), r = IDX(m, i, j + 1, k);
                        double fl = 0.5 * (ri + rho[l]), fr = 0.5 * (ri + rho[r]);
                        double d2 = 1.0 / (DY(m, j) * DY(m, j));
                        diag += alpha * d2 * (1.0 / fl + 1.0 / fr);
                        off += alpha * d2 * (sg[l] / fl + sg[r] / fr);
                    }
                    if (dim >= 3) {
                        size_t l = IDX(m, i, j, k - 1), r = IDX(m, i, j, k + 1);
                        double fl = 0.5 * (ri + rho[l]), fr = 0.5 * (ri + rho[r]);
                        double d2 = 1.0 / (DZ(m, k) * DZ(m, k));
                        diag += alpha * d2 * (1.0 / fl + 1.0 / fr);
                        off += alpha * d2 * (sg[l] / fl + sg[r] / fr);
                    }
                    sg[id] = (rhs + off) / diag;
                }
        fill_scalar(m, sg);
    }
}

void or_igr(or_solver* s, double* const* f, const sifv_config* c) { igr_solve(s, f, c); }

/* ------------------------------------------------------------------ */
/* time step: src/RKTimeStepping.cpp:15-123                             */
/* ------------------------------------------------------------------ */
double or_acoustic_dt(or_solver* s, double* const* f, const sifv_config* c, double cfl,
                      double maxDt) {
    or_mesh* m = &s->m;
    int nph = n_phases(c);
    double dt = maxDt;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                double cs;
                if (nph > 0) {
                    double a[MAXPH], ar[MAXPH];
                    for (int ph = 0; ph < nph; ++ph) a[ph] = f[SIFV_F_ALPHA + ph][id];
                    for (int ph = 0; ph < nph; ++ph) ar[ph] = f[SIFV_F_ALPHARHO + ph][id];
                    cs = mix_sound(f[SIFV_F_RHO][id], f[SIFV_F_PRES][id], a, ar, nph, c);
                } else {
                    cs = sqrt(s->gamma * dmax(f[SIFV_F_PRES][id] + s->pInf, 1e-14) /
                              dmax(f[SIFV_F_RHO][id], 1e-14));
                }
                double dc = DX(m, i) / (fabs(f[SIFV_F_VELU][id]) + cs);
                if (m->dim >= 2) dc = dmin(dc, DY(m, j) / (fabs(f[SIFV_F_VELV][id]) + cs));
                if (m->dim >= 3) dc = dmin(dc, DZ(m, k) / (fabs(f[SIFV_F_VELW][id]) + cs));
                dt = dmin(dt, cfl * dc);
            }
    return dt;
}

double or_advective_dt(or_solver* s, double* const* f, double cfl, double maxDt) {
    or_mesh* m = &s->m;
    double dt = maxDt;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                size_t id = IDX(m, i, j, k);
                double dc = DX(m, i) / dmax(fabs(f[SIFV_F_VELU][id]), 1e-14);
                if (m->dim >= 2) dc = dmin(dc, DY(m, j) / dmax(fabs(f[SIFV_F_VELV][id]), 1e-14));
                if (m->dim >= 3) dc = dmin(dc, DZ(m, k) / dmax(fabs(f[SIFV_F_VELW][id]), 1e-14));
                dt = dmin(dt, cfl * dc);
            }
    return dt;
}

/* ------------------------------------------------------------------ */
/* explicit step: ExplicitSolver::step, src/ExplicitSolver.cpp:68-203   */
/* ------------------------------------------------------------------ */
static const double RK[3][3][4] = {
    {{1, 0, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}},
    {{1, 0, 1, 1}, {1, 1, 1, 2}, {0, 0, 0, 0}},
    {{1, 0, 1, 1}, {1, 3, 1, 4}, {2, 1, 2, 3}},
};

static int cons_ids(const sifv_config* c, int dim, int* ids) {
    int n = 0;
    ids[n++] = SIFV_F_RHO;
    ids[n++] = SIFV_F_RHOU;
    if (dim >= 2) ids[n++] = SIFV_F_RHOV;
    if (dim >= 3) ids[n++] = SIFV_F_RHOW;
    ids[n++] = SIFV_F_RHOE;
    for (int ph = 0; ph < n_phases(c); ++ph) ids[n++] = SIFV_F_ALPHARHO + ph;
    for (int ph = 0; ph < n_phases(c); ++ph) ids[n++] = SIFV_F_ALPHA + ph;
    return n;
}

/* multiphase post-update clamp/normalise, src/ExplicitSolver.cpp:169-185 */
static void mp_post(double* const* f, size_t id, int nph, double aMin) {
    double rs = 0.0;
    for (int ph = 0; ph < nph; ++ph) {
        f[SIFV_F_ALPHARHO + ph][id] = dmax(f[SIFV_F_ALPHARHO + ph][id], 1e-14);
        rs += f[SIFV_F_ALPHARHO + ph][id];
    }
    f[SIFV_F_RHO][id] = rs;
    for (int ph = 0; ph < nph; ++ph) f[SIFV_F_ALPHA + ph][id] = dmax(f[SIFV_F_ALPHA + ph][id], aMin);
    double as = 0.0;
    for (int ph = 0; ph < nph; ++ph) as += f[SIFV_F_ALPHA + ph][id];
    for (int ph = 0; ph < nph; ++ph) f[SIFV_F_ALPHA + ph][id] /= as;
}

double or_explicit_step(or_solver* s, double* const* f, const sifv_config* c, double targetDt) {
    or_mesh* m = &s->m;
    double dt;
    if (c->explicitParams.constDt > 0)
        dt = c->explicitParams.constDt;
    else
    {
        dt = or_acoustic_dt(s, f, c, c->explicitParams.cfl, c->explicitParams.maxDt);
        /* src/ExplicitSolver.cpp:78-81 */
        if (has_viscosity(c))
            dt = dmin(dt, viscous_dt(s, f, c->viscousMu, c->explicitParams.cfl, c->explicitParams.maxDt));
        /* src/ExplicitSolver.cpp:82-85 */
        if (c->surfaceTensionSigma > 0.0)
            dt = dmin(dt, capillary_dt(s, f, c->surfaceTensionSigma, c->explicitParams.cfl,
                                       c->explicitParams.maxDt));
    }
    if (targetDt > 0) dt = dmin(dt, targetDt);
    int ord = c->RKOrder;
    int nph = n_phases(c);
    int ids[5 + 2 * MAXPH];
    int nc = cons_ids(c, m->dim, ids);
    for (int st = 0; st < ord; ++st) {
        cons_to_prim(s, f, c);
        fill_ghosts(m, f, SIFV_VARSET_PRIM);
        if (c->useIGR) igr_solve(s, f, c);
        compute_rhs(s, f, c, 0);
        double c1 = RK[ord - 1][st][0], c2 = RK[ord - 1][st][1];
        double c3 = RK[ord - 1][st][2], c4 = RK[ord - 1][st][3];
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    if (st == 0 && ord > 1) {
                        /* saveConservativeCell, src/SolutionState.cpp:228-240 */
                        for (int q = 0; q < nc; ++q) s->bak[ids[q]][id] = f[ids[q]][id];
                        s->bak[SIFV_F_PRES][id] = f[SIFV_F_PRES][id];
                    }
                    for (int q = 0; q < nc; ++q) {
                        double* u = f[ids[q]];
                        double* r = s->rhs[ids[q]];
                        if (ord == 1)
                            u[id] += dt * r[id];
                        else
                            u[id] = (c1 * u[id] + c2 * s->bak[ids[q]][id] + c3 * dt * r[id]) / c4;
                    }
                    if (nph > 0) mp_post(f, id, nph, c->multiPhaseParams.alphaMin);
                }
    }
    cons_to_prim(s, f, c);
    fill_ghosts(m, f, SIFV_VARSET_PRIM);
    return dt;
}

void or_compute_rhs(or_solver* s, double* const* f, const sifv_config* c, double* const* rhs) {
    compute_rhs(s, f, c, c->semiImplicit);
    int ids[5 + 2 * MAXPH];
    int nc = cons_ids(c, s->m.dim, ids);
    for (int q = 0; q < nc; ++q)
        if (rhs[ids[q]]) memcpy(rhs[ids[q]], s->rhs[ids[q]], s->m.total * sizeof(double));
    if (rhs[SIFV_F_PRES]) memcpy(rhs[SIFV_F_PRES], s->rhsP, s->m.total * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* Jacobi: src/JacobiPressureSolver.cpp:57-108 + src/PressureLaplacian.cpp:6-60 */
/* ------------------------------------------------------------------ */
static double laplacian(const or_mesh* m, const double* rho, const double* p, int i, int j, int k,
                        double* off) {
    *off = 0.0;
    double diag = 0.0;
    size_t id = IDX(m, i, j, k);
    {
        size_t a = IDX(m, i - 1, j, k), b = IDX(m, i + 1, j, k);
        double rL = 0.5 * (rho[id] + rho[a]), rR = 0.5 * (rho[id] + rho[b]);
        double dL = 0.5 * (DX(m, i - 1) + DX(m, i)), dR = 0.5 * (DX(m, i) + DX(m, i + 1));
        double cL = 1.0 / (dmax(rL, 1e-14) * dL * DX(m, i));
        double cR = 1.0 / (dmax(rR, 1e-14) * dR * DX(m, i));
        *off += cL * p[a] + cR * p[b];
        diag += cL + cR;
    }
    if (m->dim >= 2) {
        size_t a = IDX(m, i, j - 1, k), b = IDX(m, i, j + 1, k);
        double rL = 0.5 * (rho[id] + rho[a]), rR = 0.5 * (rho[id] + rho[b]);
        double dL = 0.5 * (DY(m, j - 1) + DY(m, j)), dR = 0.5 * (DY(m, j) + DY(m, j + 1));
        double cL = 1.0 / (dmax(rL, 1e-14) * dL * DY(m, j));
        double cR = 1.0 / (dmax(rR, 1e-14) * dR * DY(m, j));
        *off += cL * p[a] + cR * p[b];
        diag += cL + cR;
    }
    if (m->dim >= 3) {
        size_t a = IDX(m, i, j, k - 1), b = IDX(m, i, j, k + 1);
        double rL = 0.5 * (rho[id] + rho[a]), rR = 0.5 * (rho[id] + rho[b]);
        double dL = 0.5 * (DZ(m, k - 1) + DZ(m, k)), dR = 0.5 * (DZ(m, k) + DZ(m, k + 1));
        double cL = 1.0 / (dmax(rL, 1e-14) * dL * DZ(m, k));
        double cR = 1.0 / (dmax(rR, 1e-14) * dR * DZ(m, k));
        *off += cL * p[a] + cR * p[b];
        diag += cL + cR;
    }
    return diag;
}

int or_jacobi(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* p,
              double dt, double tol, int maxIter, double* hist) {
    or_mesh* m = &s->m;
    double dt2 = dt * dt;
    double* pn = s->pNew;
    memset(pn, 0, m->total * sizeof(double));
    for (int it = 0; it < maxIter; ++it) {
        fill_scalar(m, p);
        double mr = 0.0;
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double co = rhoc2[id] * dt2;
                    double off;
                    double dg = laplacian(m, rho, p, i, j, k, &off);
                    double den = 1.0 + co * dg;
                    pn[id] = (rhs[id] + co * off) / den;
                    mr = dmax(mr, fabs(pn[id] - p[id]));
                }
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) p[IDX(m, i, j, k)] = pn[IDX(m, i, j, k)];
        if (hist) hist[it] = mr;
        if (mr < tol) return it + 1;
    }
    return maxIter;
}

/* ------------------------------------------------------------------ */
/* PCG pressure solve.  The system is the one every PressureSolver solves:
 * (1 + c L) p = rhs with c = rhoc2 dt^2 and L the variable-coefficient
 * Laplacian of pressureLaplacian (src/PressureLaplacian.cpp:6-60), with the
 * face coefficients of non-periodic domain faces zeroed as
 * HyprePressureSolver::assembleSystem does (src/HyprePressureSolver.cpp:161-290;
 * the zero-gradient scalar ghosts of the Jacobi and Gauss-Seidel solvers give
 * the same rows).  Each row scaled by V/c (V the cell volume):
 *   S_ii = V/c + sum_f w_f,  S_i,nb(f) = -w_f,
 *   w_f = area_f / (max(0.5 (rho_i + rho_nb), 1e-14) * 0.5 (h_i + h_nb)),
 * which is V times the reference's row coefficient cL / cR up to rounding.
 * Both rows of an interior face compute the same w_f, so S is symmetric
 * (positive definite).  A periodic axis whose first ghost widths differ from
 * the wrapped cells' widths beyond rounding (a stretched periodic axis: the mesh mirrors the
 * boundary widths into the ghosts) leaves the two rows of the wrap face
 * unequal, as in the reference's own matrix; conjugate gradients need the
 * symmetric form, so such meshes are refused (return -1).  Conjugate gradients on S p = (V/c) rhs with the
 * diagonal of S as preconditioner (the HYPRE_StructPCG role,
 * src/HyprePressureSolver.cpp:119-138, with a diagonal instead of a PFMG
 * V-cycle); stop when ||rhs - (1 + c L) p||_2 <= tol ||rhs||_2 (two-norm,
 * relative to the right-hand side: src/HyprePressureSolver.cpp:350-355),
 * compared as squares.  Returns the iteration count, 0 when the initial
 * guess already passes, maxIter when not converged (or when p.Sp stops being
 * positive); rhs = 0 gives p = 0.
 * Only interior cells of p are written.
 * Sums run in the interior loop order (k, j, i). */
/* ------------------------------------------------------------------ */
static double pcg_row_w(const or_mesh* m, const double* rho, int dir, int side, int i, int j, int k) {
    /* weight of the low (side 0) or high (side 1) face of cell (i,j,k) in the
     * cell's own row, with the cell's own operand order as the reference's row */
    const int n = dir == 0 ? m->nx : dir == 1 ? m->ny : m->nz;
    const int c = dir == 0 ? i : dir == 1 ? j : k;
    if (side == 0 && c == 0 && !m->wrap[2 * dir]) return 0.0;
    if (side == 1 && c == n - 1 && !m->wrap[2 * dir + 1]) return 0.0;
    const int sg = side ? 1 : -1;
    const int ei = dir == 0 ? sg : 0, ej = dir == 1 ? sg : 0, ek = dir == 2 ? sg : 0;
    const double rf = 0.5 * (rho[IDX(m, i, j, k)] + rho[IDX(m, i + ei, j + ej, k + ek)]);
    double hf;
    if (dir == 0) hf = side ? 0.5 * (DX(m, i) + DX(m, i + 1)) : 0.5 * (DX(m, i - 1) + DX(m, i));
    else if (dir == 1) hf = side ? 0.5 * (DY(m, j) + DY(m, j + 1)) : 0.5 * (DY(m, j - 1) + DY(m, j));
    else hf = side ? 0.5 * (DZ(m, k) + DZ(m, k + 1)) : 0.5 * (DZ(m, k - 1) + DZ(m, k));
    const double area = dir == 0 ? DY(m, j) * DZ(m, k) : dir == 1 ? DX(m, i) * DZ(m, k) : DX(m, i) * DY(m, j);
    return area / (dmax(rf, 1e-14) * hf);
}

/* out = S v over the interior (v's ghosts filled by the caller), as
 * (V/c) v_i + sum_f w_f (v_i - v_nb): the differences of neighbouring values
 * are formed first, so a smooth v (a pressure ~1e5 with small variations)
 * loses nothing to the cancellation that d v_i - sum w v_nb would suffer when
 * the face terms outweigh V/c by many orders (fine meshes, large dt). */
static void pcg_apply(const or_mesh* m, const double* rho, const double* sc, const double* v, double* out) {
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                const double vi = v[id];
                double acc = sc[id] * vi;
                for (int dir = 0; dir < m->dim; ++dir) {
                    const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
                    acc += pcg_row_w(m, rho, dir, 0, i, j, k) * (vi - v[IDX(m, i - ei, j - ej, k - ek)]);
                    acc += pcg_row_w(m, rho, dir, 1, i, j, k) * (vi - v[IDX(m, i + ei, j + ej, k + ek)]);
                }
                out[id] = acc;
            }
}

static int or_pcg_symmetric(const or_mesh* m) {
    for (int dir = 0; dir < m->dim; ++dir) {
        if (!m->wrap[2 * dir]) continue;
        const int n = dir == 0 ? m->nx : dir == 1 ? m->ny : m->nz;
        double lo, last, hi, first;
        if (dir == 0) { lo = DX(m, -1); last = DX(m, n - 1); hi = DX(m, n); first = DX(m, 0); }
        else if (dir == 1) { lo = DY(m, -1); last = DY(m, n - 1); hi = DY(m, n); first = DY(m, 0); }
        else { lo = DZ(m, -1); last = DZ(m, n - 1); hi = DZ(m, n); first = DZ(m, 0); }
        /* widths equal up to the rounding of the node coordinates */
        if (fabs(lo - last) > 1e-10 * last || fabs(hi - first) > 1e-10 * first) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* Geometric multigrid V-cycle used as the PCG preconditioner (the role of
 * the PFMG preconditioner the reference's HyprePressureSolver attaches to
 * StructPCG: one V-cycle, weighted-Jacobi relaxation, one pre- and one
 * post-sweep, src/HyprePressureSolver.cpp:123-138).  Not PFMG itself: full
 * 2:1 coarsening of every active axis with piecewise-constant transfers,
 * coarse operators by rediscretisation.
 *   coarse cell C = children {2c, 2c+1} per axis (those inside the finer grid);
 *   s_C = sum of the children's s (the mass term V/c adds up exactly);
 *   w_d(C) = f_d * sum of the finer low-face weights along d of the children
 *            on C's low side, f_d = 1/2 when the axis was coarsened (the face
 *            distance doubles), 1 when it has one cell;
 *   d_C = s_C + sum_d (w_d(C) + w_d(C + e_d, wrapped)).
 * Levels are added while the finest has no coarse level yet or the coarsest
 * still has an axis longer than 4 cells.  V-cycle on level l with right-hand
 * side b into out (zero initial guess):
 *   x = om b / d;  b_{l+1}(C) = sum_children (b - A x);  V-cycle(l+1) -> t_{l+1};
 *   x += t_{l+1}(parent);  out = x + om (b - A x) / d;
 * on the coarsest level MG_COARSE_SWEEPS weighted-Jacobi sweeps from zero.
 * om = 2/3 (1D), 4/5 (2D), 6/7 (3D).  Every step is linear and symmetric in
 * the S-inner-product sense (Jacobi pre = post, restriction = transpose of
 * the prolongation), so the preconditioner is a fixed symmetric operator.
 * Children and sums run in (k, j, i) order. */
#define MG_COARSE_SWEEPS 16
#define MG_MAX_LEVELS 24

static double mg_omega(int dim) { return dim == 1 ? 2.0 / 3.0 : dim == 2 ? 0.8 : 6.0 / 7.0; }

static size_t mg_idx(const struct or_mglev* L, int i, int j, int k) {
    return (size_t)i + (size_t)L->n[0] * ((size_t)j + (size_t)L->n[1] * (size_t)k);
}

/* (A v)_C on a coarse level */
static double mg_row(const struct or_mglev* L, int dim, const double* v, int i, int j, int k) {
    const size_t id = mg_idx(L, i, j, k);
    const double vi = v[id];
    double acc = L->s[id] * vi;
    for (int dir = 0; dir < dim; ++dir) {
        const int n = L->n[dir];
        const int c = dir == 0 ? i : dir == 1 ? j : k;
        const int cm = (c - 1 + n) % n, cp = (c + 1) % n;
        const size_t lo = dir == 0 ? mg_idx(L, cm, j, k) : dir == 1 ? mg_idx(L, i, cm, k) : mg_idx(L, i, j, cm);
        const size_t hi = dir == 0 ? mg_idx(L, cp, j, k) : dir == 1 ? mg_idx(L, i, cp, k) : mg_idx(L, i, j, cp);
        acc += L->w[dir][id] * (vi - v[lo]);
        acc += L->w[dir][hi] * (vi - v[hi]);
    }
    return acc;
}

/* Builds (once per mesh) the level dimensions and storage. */
static void mg_alloc(or_solver* s) {
    if (s->mgLevels) return;
    const or_mesh* m = &s->m;
    int n[3] = {m->nx, m->ny, m->nz};
    struct or_mglev tmp[MG_MAX_LEVELS];
    int L = 0;
    while (L < MG_MAX_LEVELS) {
        const int mx = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
        if (L > 0 && mx <= 4) break;
        for (int d = 0; d < 3; ++d) n[d] = (n[d] + 1) / 2;
        struct or_mglev* c = &tmp[L];
        memset(c, 0, sizeof *c);
        for (int d = 0; d < 3; ++d) c->n[d] = n[d];
        c->cells = (size_t)n[0] * n[1] * n[2];
        c->s = zalloc(c->cells);
        for (int d = 0; d < 3; ++d) c->w[d] = zalloc(c->cells);
        c->d = zalloc(c->cells);
        c->b = zalloc(c->cells);
        c->x = zalloc(c->cells);
        c->t = zalloc(c->cells);
        ++L;
    }
    s->mg = (struct or_mglev*)malloc(sizeof(struct or_mglev) * L);
    memcpy(s->mg, tmp, sizeof(struct or_mglev) * L);
    s->mgLevels = L;
}

/* Coarse operators for this solve: level 1 from the fine s and face weights
 * (rho), deeper levels from the level above. */
static void mg_setup(or_solver* s, const double* rho, const double* sc) {
    const or_mesh* m = &s->m;
    const int dim = m->dim;
    for (int l = 0; l < s->mgLevels; ++l) {
        struct or_mglev* C = &s->mg[l];
        const struct or_mglev* F = l ? &s->mg[l - 1] : NULL;
        const int nf[3] = {F ? F->n[0] : m->nx, F ? F->n[1] : m->ny, F ? F->n[2] : m->nz};
        double fac[3];
        for (int d = 0; d < 3; ++d) fac[d] = nf[d] != C->n[d] ? 0.5 : 1.0;
        for (int ck = 0; ck < C->n[2]; ++ck)
            for (int cj = 0; cj < C->n[1]; ++cj)
                for (int ci = 0; ci < C->n[0]; ++ci) {
                    double ss = 0.0, ww[3] = {0.0, 0.0, 0.0};
                    for (int k = 2 * ck; k < 2 * ck + 2 && k < nf[2]; ++k)
                        for (int j = 2 * cj; j < 2 * cj + 2 && j < nf[1]; ++j)
                            for (int i = 2 * ci; i < 2 * ci + 2 && i < nf[0]; ++i) {
                                const int at[3] = {i == 2 * ci, j == 2 * cj, k == 2 * ck};
                                if (F) {
                                    const size_t f = mg_idx(F, i, j, k);
                                    ss += F->s[f];
                                    for (int d = 0; d < dim; ++d)
                                        if (at[d]) ww[d] += F->w[d][f];
                                } else {
                                    ss += sc[IDX(m, i, j, k)];
                                    for (int d = 0; d < dim; ++d)
                                        if (at[d]) ww[d] += pcg_row_w(m, rho, d, 0, i, j, k);
                                }
                            }
                    const size_t c = mg_idx(C, ci, cj, ck);
                    C->s[c] = ss;
                    for (int d = 0; d < 3; ++d) C->w[d][c] = d < dim ? fac[d] * ww[d] : 0.0;
                }
        for (int ck = 0; ck < C->n[2]; ++ck)
            for (int cj = 0; cj < C->n[1]; ++cj)
                for (int ci = 0; ci < C->n[0]; ++ci) {
                    const size_t c = mg_idx(C, ci, cj, ck);
                    double dd = C->s[c];
                    for (int d = 0; d < dim; ++d) {
                        const int cc[3] = {ci, cj, ck};
                        int up[3] = {ci, cj, ck};
                        up[d] = (cc[d] + 1) % C->n[d];
                        dd += C->w[d][c];
                        dd += C->w[d][mg_idx(C, up[0], up[1], up[2])];
                    }
                    C->d[c] = dd;
                }
    }
}

/* out = src + om (b - A src) / d on a coarse level */
static void mg_smooth(const struct or_mglev* L, int dim, double om, const double* src, double* out) {
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                const size_t id = mg_idx(L, i, j, k);
                out[id] = src[id] + om * (L->b[id] - mg_row(L, dim, src, i, j, k)) / L->d[id];
            }
}

/* V-cycle on coarse level l (right-hand side L->b) into L->t */
static void mg_vcycle(or_solver* s, int l) {
    const int dim = s->m.dim;
    const double om = mg_omega(dim);
    struct or_mglev* L = &s->mg[l];
    for (size_t c = 0; c < L->cells; ++c) L->x[c] = om * L->b[c] / L->d[c];
    if (l == s->mgLevels - 1) {
        for (int sw = 1; sw < MG_COARSE_SWEEPS; ++sw) {
            if (sw & 1) mg_smooth(L, dim, om, L->x, L->t);
            else mg_smooth(L, dim, om, L->t, L->x);
        }
        return;
    }
    struct or_mglev* C = &s->mg[l + 1];
    for (int ck = 0; ck < C->n[2]; ++ck)
        for (int cj = 0; cj < C->n[1]; ++cj)
            for (int ci = 0; ci < C->n[0]; ++ci) {
                double acc = 0.0;
                for (int k = 2 * ck; k < 2 * ck + 2 && k < L->n[2]; ++k)
                    for (int j = 2 * cj; j < 2 * cj + 2 && j < L->n[1]; ++j)
                        for (int i = 2 * ci; i < 2 * ci + 2 && i < L->n[0]; ++i)
                            acc += L->b[mg_idx(L, i, j, k)] - mg_row(L, dim, L->x, i, j, k);
                C->b[mg_idx(C, ci, cj, ck)] = acc;
            }
    mg_vcycle(s, l + 1);
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) L->x[mg_idx(L, i, j, k)] += C->t[mg_idx(C, i / 2, j / 2, k / 2)];
    mg_smooth(L, dim, om, L->x, L->t);
}

/* z = V-cycle(r) on the fine level; q is scratch (the fine iterate).  Fine
 * products read q through its ghost fill. */
static void mg_apply(or_solver* s, const double* rho, const double* sc, const double* d, const double* r, double* q,
                     double* z) {
    const or_mesh* m = &s->m;
    const int dim = m->dim;
    const double om = mg_omega(dim);
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                q[id] = om * r[id] / d[id];
            }
    fill_scalar(m, q);
    struct or_mglev* C = &s->mg[0];
    for (int ck = 0; ck < C->n[2]; ++ck)
        for (int cj = 0; cj < C->n[1]; ++cj)
            for (int ci = 0; ci < C->n[0]; ++ci) {
                double acc = 0.0;
                for (int k = 2 * ck; k < 2 * ck + 2 && k < m->nz; ++k)
                    for (int j = 2 * cj; j < 2 * cj + 2 && j < m->ny; ++j)
                        for (int i = 2 * ci; i < 2 * ci + 2 && i < m->nx; ++i) {
                            const size_t id = IDX(m, i, j, k);
                            const double vi = q[id];
                            double a = sc[id] * vi;
                            for (int dir = 0; dir < dim; ++dir) {
                                const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
                                a += pcg_row_w(m, rho, dir, 0, i, j, k) * (vi - q[IDX(m, i - ei, j - ej, k - ek)]);
                                a += pcg_row_w(m, rho, dir, 1, i, j, k) * (vi - q[IDX(m, i + ei, j + ej, k + ek)]);
                            }
                            acc += r[id] - a;
                        }
                C->b[mg_idx(C, ci, cj, ck)] = acc;
            }
    mg_vcycle(s, 0);
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) q[IDX(m, i, j, k)] += C->t[mg_idx(C, i / 2, j / 2, k / 2)];
    fill_scalar(m, q);
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                const double vi = q[id];
                double a = sc[id] * vi;
                for (int dir = 0; dir < dim; ++dir) {
                    const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
                    a += pcg_row_w(m, rho, dir, 0, i, j, k) * (vi - q[IDX(m, i - ei, j - ej, k - ek)]);
                    a += pcg_row_w(m, rho, dir, 1, i, j, k) * (vi - q[IDX(m, i + ei, j + ej, k + ek)]);
                }
                z[id] = vi + om * (r[id] - a) / d[id];
            }
}

static int pcg_core(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* x, double dt,
                    double tol, int maxIter, int mg) {
    or_mesh* m = &s->m;
    if (!or_pcg_symmetric(m)) {
        snprintf(g_err, sizeof g_err, "PCG needs a symmetric system: a periodic axis with stretched wrap widths");
        return -1;
    }
    for (int q = 0; q < 6; ++q)
        if (!s->cg[q]) s->cg[q] = zalloc(m->total);
    double *r = s->cg[0], *z = s->cg[1], *p = s->cg[2], *q = s->cg[3], *d = s->cg[4], *sc = s->cg[5];
    const double dt2 = dt * dt;
    const double tol2 = tol * tol;
    double bb = 0.0;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                sc[id] = VOL(m, i, j, k) / (rhoc2[id] * dt2);
                double dd = sc[id];
                for (int dir = 0; dir < m->dim; ++dir) {
                    dd += pcg_row_w(m, rho, dir, 0, i, j, k);
                    dd += pcg_row_w(m, rho, dir, 1, i, j, k);
                }
                d[id] = dd;
                bb += rhs[id] * rhs[id];
            }
    if (bb == 0.0) {
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) x[IDX(m, i, j, k)] = 0.0;
        return 0;
    }
    /* the caller's ghosts stay as they are: the matrix product reads a copy */
    memcpy(z, x, m->total * sizeof(double));
    fill_scalar(m, z);
    pcg_apply(m, rho, sc, z, q);
    double rr = 0.0, rz = 0.0;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                r[id] = sc[id] * rhs[id] - q[id];
                const double ro = r[id] / sc[id];
                rr += ro * ro;
                z[id] = r[id] / d[id];
                rz += r[id] * z[id];
                p[id] = z[id];
            }
    if (mg) {
        mg_alloc(s);
        mg_setup(s, rho, sc);
        mg_apply(s, rho, sc, d, r, q, z);
        rz = 0.0;
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    const size_t id = IDX(m, i, j, k);
                    rz += r[id] * z[id];
                    p[id] = z[id];
                }
    }
    if (rr <= tol2 * bb) return 0;
    for (int it = 0; it < maxIter; ++it) {
        fill_scalar(m, p);
        pcg_apply(m, rho, sc, p, q);
        double pq = 0.0;
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    const size_t id = IDX(m, i, j, k);
                    pq += p[id] * q[id];
                }
        /* breakdown (p.Sp not positive: rounding has exhausted the search
         * directions): stop, reported as not converged */
        if (!(pq > 0.0)) return maxIter;
        const double alpha = rz / pq;
        double rzn = 0.0;
        rr = 0.0;
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    const size_t id = IDX(m, i, j, k);
                    x[id] += alpha * p[id];
                    r[id] -= alpha * q[id];
                    const double ro = r[id] / sc[id];
                    rr += ro * ro;
                    z[id] = r[id] / d[id];
                    rzn += r[id] * z[id];
                }
        if (mg) {
            mg_apply(s, rho, sc, d, r, q, z);
            rzn = 0.0;
            for (int k = 0; k < m->nz; ++k)
                for (int j = 0; j < m->ny; ++j)
                    for (int i = 0; i < m->nx; ++i) {
                        const size_t id = IDX(m, i, j, k);
                        rzn += r[id] * z[id];
                    }
        }
        if (rr <= tol2 * bb) return it + 1;
        /* a preconditioned residual with r.z <= 0 (rounding at the end of a
         * solve): stop, reported as not converged */
        if (!(rzn > 0.0)) return maxIter;
        const double beta = rzn / rz;
        rz = rzn;
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    const size_t id = IDX(m, i, j, k);
                    p[id] = z[id] + beta * p[id];
                }
    }
    return maxIter;
}


int or_pcg(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* x, double dt,
           double tol, int maxIter) {
    return pcg_core(s, rho, rhoc2, rhs, x, dt, tol, maxIter, 0);
}

/* z = one V-cycle of the multigrid preconditioner applied to r (interior
 * cells of padded fields), for the system of or_pcg on rho / rhoc2 / dt.
 * Returns -1 where or_pcg refuses the mesh. */
int or_mg_precond(or_solver* s, const double* rho, const double* rhoc2, double dt, const double* r, double* z) {
    or_mesh* m = &s->m;
    if (!or_pcg_symmetric(m)) {
        snprintf(g_err, sizeof g_err, "PCG needs a symmetric system: a periodic axis with stretched wrap widths");
        return -1;
    }
    for (int q = 0; q < 6; ++q)
        if (!s->cg[q]) s->cg[q] = zalloc(m->total);
    double *q = s->cg[3], *d = s->cg[4], *sc = s->cg[5];
    const double dt2 = dt * dt;
    for (int k = 0; k < m->nz; ++k)
        for (int j = 0; j < m->ny; ++j)
            for (int i = 0; i < m->nx; ++i) {
                const size_t id = IDX(m, i, j, k);
                sc[id] = VOL(m, i, j, k) / (rhoc2[id] * dt2);
                double dd = sc[id];
                for (int dir = 0; dir < m->dim; ++dir) {
                    dd += pcg_row_w(m, rho, dir, 0, i, j, k);
                    dd += pcg_row_w(m, rho, dir, 1, i, j, k);
                }
                d[id] = dd;
            }
    mg_alloc(s);
    mg_setup(s, rho, sc);
    mg_apply(s, rho, sc, d, r, q, z);
    return s->mgLevels;
}

/* PCG as or_pcg with one multigrid V-cycle as the preconditioner instead of
 * the diagonal (the StructPCG + PFMG pairing of the reference's
 * HyprePressureSolver, src/HyprePressureSolver.cpp:123-138, 350-359). */
int or_mgpcg(or_solver* s, const double* rho, const double* rhoc2, const double* rhs, double* x, double dt,
             double tol, int maxIter) {
    return pcg_core(s, rho, rhoc2, rhs, x, dt, tol, maxIter, 1);
}

void or_set_pressure_solver(or_solver* s, int kind) { s->psolver = kind; }

/* ------------------------------------------------------------------ */
/* semi-implicit step: SemiImplicitSolver::step, src/SemiImplicitSolver.cpp:106-236 */
/* ------------------------------------------------------------------ */
double or_semi_step(or_solver* s, double* const* f, const sifv_config* c, double targetDt,
                    int* iters) {
    or_mesh* m = &s->m;
    int dim = m->dim;
    int nph = n_phases(c);
    const int mom[3] = {SIFV_F_RHOU, SIFV_F_RHOV, SIFV_F_RHOW};
    const int vel[3] = {SIFV_F_VELU, SIFV_F_VELV, SIFV_F_VELW};
    double dt = or_advective_dt(s, f, c->semiImplicitParams.cfl, c->semiImplicitParams.maxDt);
    /* src/SemiImplicitSolver.cpp:112-115 */
    if (has_viscosity(c))
        dt = dmin(dt, viscous_dt(s, f, c->viscousMu, c->semiImplicitParams.cfl, c->semiImplicitParams.maxDt));
    if (c->surfaceTensionSigma > 0.0)
        dt = dmin(dt, capillary_dt(s, f, c->surfaceTensionSigma, c->semiImplicitParams.cfl,
                                   c->semiImplicitParams.maxDt));
    if (targetDt > 0) dt = dmin(dt, targetDt);
    /* std::clamp(dt, minDt, maxDt) */
    if (dt < c->semiImplicitParams.minDt) dt = c->semiImplicitParams.minDt;
    else if (c->semiImplicitParams.maxDt < dt) dt = c->semiImplicitParams.maxDt;
    int ord = c->RKOrder;
    int lastIters = 0;
    for (int st = 0; st < ord; ++st) {
        cons_to_prim(s, f, c);
        fill_ghosts(m, f, SIFV_VARSET_PRIM);
        if (c->useIGR) igr_solve(s, f, c);
        compute_rhs(s, f, c, 1);
        double c1 = RK[ord - 1][st][0], c2 = RK[ord - 1][st][1];
        double c3 = RK[ord - 1][st][2], c4 = RK[ord - 1][st][3];
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    if (st == 0 && ord > 1) {
                        s->bak[SIFV_F_RHO][id] = f[SIFV_F_RHO][id];
                        for (int d = 0; d < dim; ++d) s->bak[mom[d]][id] = f[mom[d]][id];
                        s->bak[SIFV_F_RHOE][id] = f[SIFV_F_RHOE][id];
                        s->bak[SIFV_F_PRES][id] = f[SIFV_F_PRES][id];
                        for (int ph = 0; ph < nph; ++ph) {
                            s->bak[SIFV_F_ALPHARHO + ph][id] = f[SIFV_F_ALPHARHO + ph][id];
                            s->bak[SIFV_F_ALPHA + ph][id] = f[SIFV_F_ALPHA + ph][id];
                        }
                    }
                    s->pAdv[id] = f[SIFV_F_PRES][id];
                    double r0 = ord > 1 ? s->bak[SIFV_F_RHO][id] : 0.0;
                    double e0 = ord > 1 ? s->bak[SIFV_F_RHOE][id] : 0.0;
                    f[SIFV_F_RHO][id] =
                        (c1 * f[SIFV_F_RHO][id] + c2 * r0 + c3 * dt * s->rhs[SIFV_F_RHO][id]) / c4;
                    for (int d = 0; d < dim; ++d) {
                        double u0 = ord > 1 ? s->bak[mom[d]][id] : 0.0;
                        s->rhoUStar[d][id] =
                            (c1 * f[mom[d]][id] + c2 * u0 + c3 * dt * s->rhs[mom[d]][id]) / c4;
                    }
                    s->rhoEstar[id] =
                        (c1 * f[SIFV_F_RHOE][id] + c2 * e0 + c3 * dt * s->rhs[SIFV_F_RHOE][id]) / c4;
                    double p0 = ord > 1 ? s->bak[SIFV_F_PRES][id] : 0.0;
                    s->pAdv[id] = (c1 * s->pAdv[id] + c2 * p0 + c3 * dt * s->rhsP[id]) / c4;
                    if (nph > 0) {
                        for (int q = 0; q < 2; ++q)
                            for (int ph = 0; ph < nph; ++ph) {
                                int fid = (q ? SIFV_F_ALPHA : SIFV_F_ALPHARHO) + ph;
                                if (ord == 1)
                                    f[fid][id] += dt * s->rhs[fid][id];
                                else
                                    f[fid][id] = (c1 * f[fid][id] + c2 * s->bak[fid][id] +
                                                  c3 * dt * s->rhs[fid][id]) / c4;
                            }
                        mp_post(f, id, nph, c->multiPhaseParams.alphaMin);
                    }
                    double rs = dmax(f[SIFV_F_RHO][id], 1e-14);
                    for (int d = 0; d < dim; ++d) f[vel[d]][id] = s->rhoUStar[d][id] / rs;
                }
        fill_ghosts(m, f, SIFV_VARSET_PRIM);
        /* solvePressure, src/SemiImplicitSolver.cpp:533-588 */
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double cs;
                    if (nph > 0) {
                        double a[MAXPH], ar[MAXPH];
                        for (int ph = 0; ph < nph; ++ph) a[ph] = f[SIFV_F_ALPHA + ph][id];
                        for (int ph = 0; ph < nph; ++ph) ar[ph] = f[SIFV_F_ALPHARHO + ph][id];
                        cs = mix_sound(f[SIFV_F_RHO][id], f[SIFV_F_PRES][id], a, ar, nph, c);
                    } else {
                        cs = sqrt(s->gamma * dmax(f[SIFV_F_PRES][id] + s->pInf, 1e-14) /
                                  dmax(f[SIFV_F_RHO][id], 1e-14));
                    }
                    s->rhoc2[id] = f[SIFV_F_RHO][id] * cs * cs;
                }
        /* computeDivergence, src/SemiImplicitSolver.cpp:680-717 */
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double dv = 0.0;
                    {
                        const double* U = f[SIFV_F_VELU];
                        size_t a = IDX(m, i - 1, j, k), b = IDX(m, i + 1, j, k);
                        double fl = 0.5 * (U[a] + U[id]), fr = 0.5 * (U[id] + U[b]);
                        dv += (fr - fl) / DX(m, i);
                    }
                    if (dim >= 2) {
                        const double* V = f[SIFV_F_VELV];
                        size_t a = IDX(m, i, j - 1, k), b = IDX(m, i, j + 1, k);
                        double fl = 0.5 * (V[a] + V[id]), fr = 0.5 * (V[id] + V[b]);
                        dv += (fr - fl) / DY(m, j);
                    }
                    if (dim >= 3) {
                        const double* W = f[SIFV_F_VELW];
                        size_t a = IDX(m, i, j, k - 1), b = IDX(m, i, j, k + 1);
                        double fl = 0.5 * (W[a] + W[id]), fr = 0.5 * (W[id] + W[b]);
                        dv += (fr - fl) / DZ(m, k);
                    }
                    s->divUstar[id] = dv;
                }
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    s->pRhs[id] = s->pAdv[id] - s->rhoc2[id] * dt * s->divUstar[id];
                    s->pressure_[id] = f[SIFV_F_PRES][id];
                }
        lastIters = s->psolver >= 1
                        ? pcg_core(s, f[SIFV_F_RHO], s->rhoc2, s->pRhs, s->pressure_, dt,
                                   c->semiImplicitParams.pressureTol, c->semiImplicitParams.maxPressureIters,
                                   s->psolver == 2)
                        : or_jacobi(s, f[SIFV_F_RHO], s->rhoc2, s->pRhs, s->pressure_, dt,
                                    c->semiImplicitParams.pressureTol, c->semiImplicitParams.maxPressureIters,
                                    NULL);
        fill_scalar(m, s->pressure_);
        if (f[SIFV_F_SIGMA]) fill_scalar(m, f[SIFV_F_SIGMA]);
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    f[SIFV_F_PRES][id] = s->pressure_[id];
                }
        /* correctionStep, src/SemiImplicitSolver.cpp:590-678 */
        const double* P = s->pressure_;
        const double* S = f[SIFV_F_SIGMA];
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    for (int d = 0; d < dim; ++d) {
                        size_t a = IDX(m, i - (d == 0), j - (d == 1), k - (d == 2));
                        size_t b = IDX(m, i + (d == 0), j + (d == 1), k + (d == 2));
                        double pl = 0.5 * ((P[a] + S[a]) + (P[id] + S[id]));
                        double pr = 0.5 * ((P[id] + S[id]) + (P[b] + S[b]));
                        double h = d == 0 ? DX(m, i) : d == 1 ? DY(m, j) : DZ(m, k);
                        f[mom[d]][id] = s->rhoUStar[d][id] - dt * (pr - pl) / h;
                    }
                }
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double rs = dmax(f[SIFV_F_RHO][id], 1e-14);
                    for (int d = 0; d < dim; ++d) f[vel[d]][id] = f[mom[d]][id] / rs;
                }
        fill_ghosts(m, f, SIFV_VARSET_PRIM);
        for (int k = 0; k < m->nz; ++k)
            for (int j = 0; j < m->ny; ++j)
                for (int i = 0; i < m->nx; ++i) {
                    size_t id = IDX(m, i, j, k);
                    double rho = f[SIFV_F_RHO][id];
                    double ke = 0.5 * rho * f[vel[0]][id] * f[vel[0]][id];
                    for (int d = 1; d < dim; ++d) ke += 0.5 * rho * f[vel[d]][id] * f[vel[d]][id];
                    if (nph > 0) {
                        double a[MAXPH];
                        for (int ph = 0; ph < nph; ++ph) a[ph] = f[SIFV_F_ALPHA + ph][id];
                        f[SIFV_F_RHOE][id] = mix_energy(P[id], a, nph, ke, c);
                    } else {
                        f[SIFV_F_RHOE][id] = (P[id] + s->gamma * s->pInf) / (s->gamma - 1.0) + ke;
                    }
                }
    }
    if (iters) *iters = lastIters;
    return dt;
}
