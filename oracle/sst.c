/*
 * sst.c -- oracle restatement of the k-omega SST transport equations on hex8
 * meshes.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (file:line in the reference, src/ omitted):
 *   eddy viscosity        model/turbulence/RANS/shearStressTransport/
 *                         shearStressTransportModel.cpp:597-690
 *   effective viscosity   model/turbulence/turbulenceModel.cpp:121
 *   F1 blending           shearStressTransportModel.cpp:95-174
 *   production            shearStressTransportModel.cpp:176-340 (wall nodes:
 *                         :341-360 zero it, the wall treatment supplies it)
 *   diffusivities         equation/turbulence/RANS/shearStressTransport/
 *                         turbulentKineticEnergy/...SSTEquation.cpp setup(),
 *                         turbulentEddyFrequency/...SSTEquation.cpp:23-78
 *   k node terms          assemble/turbulence/RANS/shearStressTransport/
 *                         segregatedShearStressTransport/turbulentKineticEnergy/
 *                         turbulentKineticEnergySSTAssemblerNodeTerms.cpp:13-128
 *   omega node terms      .../turbulentEddyFrequencySSTAssemblerNodeTerms.cpp:13-219
 *   interior terms        assemble/phiAssembler/phiAssemblerElemTerms.hpp:10-413
 *   relaxation            phiAssembler.h:743-753, 966-1006; boundary rhs
 *                         relaxation :1008-1152 with 0.75
 *                         (turbulentKineticEnergyAssembler.h:44-48)
 *   omega wall rows       turbulentEddyFrequencySSTAssembler.h postAssemble_
 *   correction + clip     equation/equation.h:204-286, SSTEquation solve()
 *   solve order           segregatedShearStressTransportEquations.cpp preSolve,
 *                         solve
 *
 * Advection is upwind, with the limited deferred correction (blending
 * factors beta_k / beta_w) when the case is high-resolution.  Constant
 * density; incompressible (no dilatation terms); no frame rotation.
 */
#include "case_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* shearStressTransportModel.h:21-37, RANSModel.h:20-22, turbulenceModel.h:20 */
static const double kSigK1 = 0.85, kSigK2 = 1.0, kSigW1 = 0.5, kSigW2 = 0.856;
static const double kBeta1 = 0.075, kBeta2 = 0.0828, kGam1 = 0.5532, kGam2 = 0.4403;
static const double kA1 = 0.31, kBetaStar = 0.09, kProdLim = 10.0, kCmu = 0.09;

void or_sst_enable(or_case *c, const or_sst_params *s)
{
    const size_t nn = (size_t)c->nn;
    c->sst = *s;
    double **one[] = {&c->k, &c->w, &c->ymin, &c->mut, &c->F1, &c->Pk, &c->Pk_wall, &c->kprev};
    for (size_t i = 0; i < sizeof(one) / sizeof(one[0]); ++i)
        if (!*one[i]) *one[i] = (double *)calloc(nn, sizeof(double));
    if (!c->gradK) c->gradK = (double *)calloc(3 * nn, sizeof(double));
    if (!c->gradW) c->gradW = (double *)calloc(3 * nn, sizeof(double));
    c->k_scale = c->w_scale = 1.0;
    c->scalar_terms = 15;
}

void or_set_scalar_terms(or_case *c, int mask) { c->scalar_terms = mask; }

/* sqrt(2 S_ij S_ij), S = (grad U + grad U^T) / 2 */
static double strain_mag(const double *g)
{
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double r = 0.5 * (g[3 * i + j] + g[3 * j + i]);
            s += r * r;
        }
    return sqrt(2.0 * s);
}

void or_sst_update_mut(or_case *c)
{
    const double mu = c->prm.mu, urf = c->sst.urf_mut;
    for (int i = 0; i < c->nn; ++i) {
        const double rho = c->rho[i], k = c->k[i], w = c->w[i], y = c->ymin[i];
        const double sij = strain_mag(&c->gradU[9 * i]);
        const double minDSq = y * y;
        const double trbDiss = sqrt(k) / kBetaStar / (w + SMALL) / (y + SMALL);
        const double lamDiss = 500.0 * mu / rho / (w + SMALL) / (minDSq + SMALL);
        const double fArgTwo = fmax(2.0 * trbDiss, lamDiss);
        const double fTwo = tanh(fArgTwo * fArgTwo);
        const double mutv = kA1 * rho * k / (fmax(kA1 * w, sij * fTwo) + SMALL);
        c->mut[i] = urf * mutv + (1.0 - urf) * c->mut[i];
        c->mu[i] = mu + c->mut[i];
    }
}

void or_sst_update_F1(or_case *c)
{
    const double mu = c->prm.mu, CDkwClip = 1.0e-10;
    for (int i = 0; i < c->nn; ++i) {
        const double rho = c->rho[i], k = c->k[i], w = c->w[i], y = c->ymin[i];
        double cross = 0.0;
        for (int j = 0; j < 3; ++j) cross += c->gradK[3 * i + j] * c->gradW[3 * i + j];
        const double minDSq = y * y;
        const double trbDiss = sqrt(k) / kBetaStar / (w + SMALL) / (y + SMALL);
        const double lamDiss = 500.0 * mu / rho / (w + SMALL) / (minDSq + SMALL);
        const double CDkw = fmax(2.0 * rho * kSigW2 * cross / (w + SMALL), CDkwClip);
        const double arg = fmin(fmax(trbDiss, lamDiss),
                                4.0 * rho * kSigW2 * k / (CDkw + SMALL) / (minDSq + SMALL));
        c->F1[i] = tanh(arg * arg * arg * arg);
    }
}

/* grad U : (grad U + grad U^T), clipped at 0 */
static double pk_by_mut(const double *g)
{
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s += g[3 * i + j] * (g[3 * i + j] + g[3 * j + i]);
    return fmax(s, 0.0);
}

void or_sst_update_Pk(or_case *c)
{
    for (int i = 0; i < c->nn; ++i)
        c->Pk[i] = c->wnode[i] ? c->Pk_wall[i] : fmax(c->mut[i] * pk_by_mut(&c->gradU[9 * i]), 0.0);
}

static void node_terms(or_case *c, int which, int i, double *lhs, double *rhs)
{
    const double rho = c->rho[i], vol = c->vol[i], div = c->div[i], w = c->w[i];
    const double phi = which == 0 ? c->k[i] : w;
    double l = 0.0, r = 0.0;
    if (c->scalar_terms & 1) {
        if (c->tscheme) {  /* phiAssemblerNodeTerms.hpp:150-404 (BDF1 / BDF2) */
            double cf[3];
            or_bdf_coeff(c, cf);
            const double f0 = cf[0] * rho * vol / c->tdt;
            const double *old = which == 0 ? c->kold : c->wold;
            const double *old2 = which == 0 ? c->kold2 : c->wold2;
            l += f0;
            r -= f0 * phi + cf[1] * rho * vol / c->tdt * old[i] + cf[2] * rho * vol / c->tdt * old2[i];
        } else {
            l += rho * vol / c->prm.dt;
        }
        if (div < 0) l += -div;
        r -= -div * phi;
    }
    if (!(c->scalar_terms & 2)) {
    } else if (which == 0) {
        const double Dk = kBetaStar * rho * w * c->k[i];
        double Pk = c->Pk[i];
        if (Pk > kProdLim * Dk) Pk = kProdLim * Dk;
        r += (Pk - Dk) * vol;
        l += kBetaStar * rho * w * vol;
    } else {
        const double mu = c->prm.mu, y = c->ymin[i], kPre = c->kprev[i];
        const double *g = &c->gradU[9 * i];
        double P = pk_by_mut(g);
        const double sij = strain_mag(g);
        const double arg2 = fmin(fmax(2 * sqrt(kPre) / (kCmu * w * y + SMALL),
                                      500 * mu / (y * y * rho * w + SMALL)),
                                 100.0);
        const double F2 = tanh(arg2 * arg2);
        P = fmin(P, ((kProdLim * kBetaStar) / kA1) * w * fmax(kA1 * w, F2 * sqrt(2.0) * sij));
        double cross = 0.0;
        for (int j = 0; j < 3; ++j) cross += c->gradK[3 * i + j] * c->gradW[3 * i + j];
        const double F1 = c->F1[i], om = 1.0 - F1;
        const double beta = F1 * kBeta1 + om * kBeta2;
        const double gamma = F1 * kGam1 + om * kGam2;
        const double sigmaD = 2.0 * om * kSigW2;
        const double Pw = gamma * rho * P;
        const double Dw = beta * rho * w * w;
        const double Sw = sigmaD * rho * cross / (w + SMALL);
        r += (Pw - Dw + Sw) * vol;
        l += (2.0 * beta * rho * w + fmax(Sw / (w + SMALL), 0.0)) * vol;
    }
    *lhs = l;
    *rhs = r;
}

void or_assemble_scalar(or_case *c, int which)
{
    const int nn = c->nn;
    const or_sst_params *S = &c->sst;
    double *phiF = which == 0 ? c->k : c->w;
    const int shifted = which == 0 ? S->k_shifted : S->w_shifted;
    const int gshift = which == 0 ? S->kgrad_shifted : S->wgrad_shifted;
    const double s1 = which == 0 ? kSigK1 : kSigW1, s2 = which == 0 ? kSigK2 : kSigW2;
    memset(c->A1, 0, sizeof(double) * (size_t)c->nnz);
    memset(c->b1, 0, sizeof(double) * (size_t)nn);
    /* node terms first (phiAssembler::assemble_ :754-758) */
#pragma omp parallel for num_threads(or_nthreads) schedule(static)
    for (int i = 0; i < nn; ++i) {
        double l, r;
        node_terms(c, which, i, &l, &r);
        or_apply_coeff(1, 1, &i, &l, &r, nn, c->row_ptr, c->cols, c->diag_off, c->A1, c->b1);
    }
    /* diffusivity Gamma = mu + sigma(F1) mut (computeGamma_ / setup lambdas) */
    or_tables tab[4];
    or_tables_init(tab);
    double lhs[64], rhs[8];
    /* high-resolution advection: limited deferred correction (blending
     * factor beta_k / beta_w, phiAssemblerElemTerms.hpp:291-340) */
    const double *beta = c->prm.high_res ? (which == 0 ? c->beta_k : c->beta_w) : NULL;
    const double *gphi = which == 0 ? c->gradK : c->gradW;
    or_color_elements(c);
    if (c->scalar_terms & 4)
    OR_PARALLEL
    {
    double co[24], av[36], dndx[288], detj[12], lhs[64], rhs[8];
    OR_FOR_COLORED(c, e) {
        const int t = or_etype(c, e);
        const or_tables *T = &tab[t];
        const int npe = T->npe, *lr = T->lr;
        const double *sf = T->sf[shifted ? 1 : 0];
        const double *sfc = T->sf[0]; /* coordinates: standard shape functions */
        const int *en = &c->conn[8 * e];
        double phi[8], G[8];
        for (int a = 0; a < npe; ++a) {
            const int n = en[a];
            phi[a] = phiF[n];
            const double F1 = c->F1[n];
            G[a] = c->prm.mu + c->mut[n] * (F1 * s1 + (1.0 - F1) * s2);
        }
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        or_gather_coords(c, e, co);
        or_elem_areav(t, co, av);
        or_elem_gradop(t, gshift, co, dndx, detj);
        for (int ip = 0; ip < T->nscs; ++ip) {
            const int il = lr[2 * ip], ir = lr[2 * ip + 1];
            const double tm = c->mdot[12 * e + ip];
            const double *dn = &dndx[3 * npe * ip];
            double Gip = 0.0;
            for (int ic = 0; ic < npe; ++ic) Gip += sf[npe * ip + ic] * G[ic];
            /* upwind advection (+ deferred correction) */
            const int iu = tm > 0 ? il : ir;
            double dcorr = 0.0;
            if (beta) {
                double xip[3] = {0, 0, 0};
                for (int ic = 0; ic < npe; ++ic)
                    for (int j = 0; j < 3; ++j) xip[j] += sfc[npe * ip + ic] * co[3 * ic + j];
                for (int j = 0; j < 3; ++j)
                    dcorr += beta[en[iu]] * (xip[j] - co[3 * iu + j]) * gphi[3 * en[iu] + j];
            }
            const double aflux = tm * (phi[iu] + dcorr);
            rhs[il] -= aflux;
            rhs[ir] += aflux;
            const double aL = 0.5 * (tm + fabs(tm)), aR = 0.5 * (tm - fabs(tm));
            lhs[npe * il + il] += aL;
            lhs[npe * ir + il] -= aL;
            lhs[npe * ir + ir] -= aR;
            lhs[npe * il + ir] += aR;
            /* diffusion */
            for (int ic = 0; ic < npe; ++ic) {
                double l = 0.0;
                for (int j = 0; j < 3; ++j) l += -Gip * dn[3 * ic + j] * av[3 * ip + j];
                lhs[npe * il + ic] += l;
                lhs[npe * ir + ic] -= l;
                rhs[il] -= l * phi[ic];
                rhs[ir] += l * phi[ic];
            }
        }
        or_apply_coeff(1, npe, en, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A1, c->b1);
    }
    }
    /* inlets (fixed value), outlets and openings (zero gradient while the
     * ip is not reversed; an opening ip never is); nothing at walls (k
     * zero-gradient, omega rows fixed below) or symmetry planes */
    for (int f = 0; f < ((c->scalar_terms & 4) ? c->nbf : 0); ++f) {
        const int btype = c->bc_type[c->bf_bc[f]];
        if (btype != OR_BC_INLET && btype != OR_BC_OUTLET && btype != OR_BC_OPENING) continue;
        const int *en = &c->conn[8 * c->bf_elem[f]];
        const int npe = or_topo_npe(or_etype(c, c->bf_elem[f]));
        double G[8];
        for (int a = 0; a < npe; ++a) {
            const double F1 = c->F1[en[a]];
            G[a] = c->prm.mu + c->mut[en[a]] * (F1 * s1 + (1.0 - F1) * s2);
        }
        memset(lhs, 0, sizeof(lhs));
        memset(rhs, 0, sizeof(rhs));
        or_bc_scalar(c, which, f, G, lhs, rhs);
        or_apply_coeff(1, npe, en, lhs, rhs, nn, c->row_ptr, c->cols, c->diag_off, c->A1, c->b1);
    }
    /* postAssemble_: relaxation, boundary rhs relaxation */
    if (!(c->scalar_terms & 8)) return;
    const double uinv = 1.0 / (which == 0 ? S->urf_k : S->urf_w);
    for (int i = 0; i < nn; ++i) c->A1[c->row_ptr[i] + c->diag_off[i]] *= uinv;
    for (int i = 0; i < nn; ++i)
        if (c->bnode[i]) c->b1[i] *= 0.75;
    if (which == 1)
        for (int i = 0; i < nn; ++i)
            if (c->wnode[i]) {  /* omega rows fixed at no-slip wall nodes */
                for (int s = c->row_ptr[i]; s < c->row_ptr[i + 1]; ++s)
                    if (s - c->row_ptr[i] != c->diag_off[i]) c->A1[s] = 0.0;
                c->b1[i] = 0.0;
            }
}

void or_correct_scalar(or_case *c, int which)
{
    double *phi = which == 0 ? c->k : c->w;
    for (int i = 0; i < c->nn; ++i) {
        double v = phi[i] + c->x1[i];
        if (v < SMALL) v = SMALL;
        phi[i] = v;
    }
    or_field_scale(phi, c->nn, 1, which == 0 ? &c->k_scale : &c->w_scale);
}

void or_sst_iteration(or_case *c, const double *ctl, double *hist)
{
    double res[2], rk, rw;
    const double ks0 = c->k_scale, ws0 = c->w_scale;
    memcpy(c->kprev, c->k, sizeof(double) * (size_t)c->nn);
    or_sst_update_mut(c);
    or_sst_update_F1(c);
    or_sst_update_Pk(c);
    or_assemble_scalar(c, 0);
    or_scaled_residual(c, 1, &rk);
    memset(c->x1, 0, sizeof(double) * (size_t)c->nn);
    const int itk = or_solve(c, 1, 0, (int)ctl[0], ctl[1], 1e-30, 30, res);
    or_correct_scalar(c, 0);
    or_assemble_scalar(c, 1);
    or_scaled_residual(c, 1, &rw);
    memset(c->x1, 0, sizeof(double) * (size_t)c->nn);
    const int itw = or_solve(c, 1, 0, (int)ctl[0], ctl[1], 1e-30, 30, res);
    or_correct_scalar(c, 1);
    or_gradient(c, 2);
    or_gradient(c, 3);
    if (hist) {
        hist[0] = rk / (ks0 + SMALL);
        hist[1] = rw / (ws0 + SMALL);
        hist[2] = itk;
        hist[3] = itw;
    }
}
