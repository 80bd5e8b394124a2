/*
 * bc.c -- oracle restatement of the velocity-inlet and static-pressure-outlet
 * boundary conditions of the segregated flow and k-omega SST path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Sides of every topology: quad4 sides of hex8 / wedge6 / pyramid5 and tri3
 * sides of tet4 / wedge6 / pyramid5, one face ip per side node (side order).
 *
 * Inlet: specified velocity (per boundary id, the bc_U row), zero-gradient
 * pressure, specified k / omega.  Outlet: specified static pressure,
 * zero-gradient velocity and k / omega, with the flow-reversal flag of each
 * face ip.  Opening: specified static pressure; where the side mass flow
 * leaves, the outlet's velocity terms; where it enters, the flow is
 * entrained along a specified direction; k / omega zero-gradient; no
 * reversal flags.  Side mass flow rates live per boundary face ip ([nbf][4];
 * a tri side leaves the fourth unused).
 *
 * References (src/):
 *   side mass flow init   model/flow/flowModel.cpp:3802-4997
 *     (inlet specifiedValue: rho U_bc . A; outlet default: max(0, rho U_ip . A))
 *   side mass flow update flowModel.cpp:6470-6598 (inlet), :7323-7690 (outlet,
 *     Rhie-Chow at the face ip with the face gradient operator and the
 *     specified pressure on the face nodes)
 *   reversal flags        flowModel.cpp:15226-15454 (outlet staticPressure)
 *   mass divergence       flowModel.cpp:14237-14500 (side flows to the ip's
 *     nearest node)
 *   momentum              assemble/flow/segregatedFlow/navierStokesAssembler/
 *     navierStokesAssemblerElemBoundaryConditions.cpp:791-1124 (inlet:
 *     IP_EXPLICIT_ADVECTIVE_FLUX__, IP_FULL_STRESS_FIXED_VEL__ :344-392),
 *     :1468-1775 (outlet: IP_IMPLICIT_UPWIND_ADVECTIVE_FLUX__,
 *     IP_ZERO_NORMAL_STRESS__ :440-500); incompressible, so no divU terms
 *   pressure correction   pressureCorrectionAssembler/
 *     pressureCorrectionAssemblerElemBoundaryConditions.cpp:923-1393 (inlet),
 *     :1895-2530 (outlet)
 *   k / omega             assemble/phiAssembler/phiAssemblerElemBoundaryConditions.hpp:
 *     894-1229 (inlet fixed value), 1231-1437 (outlet zero gradient)
 *   opening               flowModel.cpp:4817-4992 (init: rho U_ip . A, not
 *     clipped), :8094-8380 (update: Rhie-Chow with the nodal pressure on
 *     every element node), :15575 (no reversal flags);
 *     navierStokesAssemblerElemBoundaryConditions.cpp:3234-3552 (msd > 0:
 *     IP_IMPLICIT_UPWIND_ADVECTIVE_FLUX__ + IP_ZERO_NORMAL_STRESS__; else
 *     IP_EXPLICIT_ADVECTIVE_FLUX__ of msd Iu d, Iu = U_ip.n / (d.n + SMALL),
 *     + IP_ZERO_NORMAL_STRESS_NO_GRADU_TRANSPOSE__ :471-519);
 *     pressureCorrectionAssemblerElemBoundaryConditions.cpp:3026-3560 (the
 *     outlet's terms at every ip, the specified pressure on the side nodes);
 *     phiAssemblerElemBoundaryConditions.hpp:1439-1760 (implicit upwind);
 *     entrainment direction velocity.cpp:4813-4970 (cartesian components,
 *     normalised, or the inward unit normal of each face ip)
 */
#include "case_internal.h"
#include "oracle.h"

#include <math.h>
#include <string.h>



static int face_type(const or_case *c, int f) { return c->bc_type[c->bf_bc[f]]; }

/* boundary types that carry side mass flow rates */
static int flow_side(int t) { return t == OR_BC_INLET || t == OR_BC_OUTLET || t == OR_BC_OPENING; }

/* entrainment direction of opening face ip `ip`: the boundary id's bc_U row
 * normalised, or (a zero row) the inward unit normal -A / |A| */
static void opening_dir(const or_case *c, int f, int ip, double d[3])
{
    const double *v = &c->bc_U[3 * c->bf_bc[f]];
    const double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (n2 > 0.0) {
        const double nv = sqrt(n2);
        for (int j = 0; j < 3; ++j) d[j] = v[j] / nv;
        return;
    }
    const double *ax = &c->bareav[12 * f + 3 * ip];
    const double am = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    for (int j = 0; j < 3; ++j) d[j] = -ax[j] / am;
}

/* the side, its face shape functions (npf x npf) and the element coordinates */
static void side_gather(const or_case *c, int f, or_side *sd, double co[24])
{
    or_side_get(c, f, sd);
    or_gather_coords(c, sd->e, co);
}

/* ------------------------------------------------------ side mass flow */
void or_bc_mdot_init(or_case *c)
{
    for (int f = 0; f < c->nbf; ++f) {
        const int t = face_type(c, f);
        if (!flow_side(t)) continue;
        or_side sd;
        double co[24], fsf[16];
        side_gather(c, f, &sd, co);
        or_side_shape(sd.npf, c->prm.u_shifted, fsf);
        const int npf = sd.npf;
        const double *Ubc = &c->bc_U[3 * c->bf_bc[f]];
        for (int ip = 0; ip < npf; ++ip) {
            const double *ax = &c->bareav[12 * f + 3 * ip];
            double rip = 0, u[3] = {0, 0, 0};
            for (int ic = 0; ic < npf; ++ic) {
                const double r = fsf[npf * ip + ic];
                rip += r * c->rho[sd.fn[ic]];
                for (int j = 0; j < 3; ++j) u[j] += r * c->U[3 * sd.fn[ic] + j];
            }
            double m = 0;
            for (int j = 0; j < 3; ++j) m += rip * (t == OR_BC_INLET ? Ubc[j] : u[j]) * ax[j];
            c->msd[4 * f + ip] = t == OR_BC_OUTLET ? fmax(0.0, m) : m;
        }
    }
}

/* Rhie-Chow mass flow at outlet / opening face ip `ip` (duRhs = du:
 * SIMPLE); the face-averaged nodal pressure gradient weighs each side node
 * 1/npf.  side_p: the specified pressure on the side's nodes (else the
 * nodal pressure field there) */
static double outlet_rhie_chow(const or_case *c, int f, int ip, const double *fdn,
                               const double *fsf, const or_side *sd, int side_p)
{
    const int npf = sd->npf, npe = sd->npe;
    const double pbc = c->bc_p[c->bf_bc[f]];
    double p8[8];
    for (int a = 0; a < npe; ++a) p8[a] = c->p[sd->en[a]];
    if (side_p)
        for (int k = 0; k < npf; ++k) p8[sd->sn[k]] = pbc;  /* nodal side pressure */
    const double fw = 1.0 / npf;
    double rip = 0, u[3] = {0, 0, 0}, du[3] = {0, 0, 0}, Gp[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    for (int ic = 0; ic < npf; ++ic) {
        const double r = fsf[npf * ip + ic];
        const int n = sd->fn[ic];
        rip += r * c->rho[n];
        for (int j = 0; j < 3; ++j) {
            u[j] += r * c->U[3 * n + j];
            du[j] += r * c->du[3 * n + j];
            Gp[j] += fw * c->gradP[3 * n + j];
        }
    }
    for (int ic = 0; ic < npe; ++ic)
        for (int j = 0; j < 3; ++j) dp[j] += fdn[3 * npe * ip + 3 * ic + j] * p8[ic];
    const double *ax = &c->bareav[12 * f + 3 * ip];
    double m = 0;
    for (int j = 0; j < 3; ++j) m += (rip * u[j] - rip * du[j] * (dp[j] - Gp[j])) * ax[j];
    return m;
}

void or_bc_mdot_update(or_case *c)
{
    const double urf = c->prm.urf_mdot;
    for (int f = 0; f < c->nbf; ++f) {
        const int t = face_type(c, f);
        if (!flow_side(t)) continue;
        or_side sd;
        double co[24], fdn[96], fdet[4], fsf[16];
        side_gather(c, f, &sd, co);
        or_side_shape(sd.npf, c->prm.u_shifted, fsf);
        const int npf = sd.npf;
        if (t != OR_BC_INLET) or_topo_side_gradop(sd.topo, c->prm.pgrad_shifted, sd.ord, co, fdn, fdet);
        const double *Ubc = &c->bc_U[3 * c->bf_bc[f]];
        for (int ip = 0; ip < npf; ++ip) {
            double *md = &c->msd[4 * f + ip];
            if (t == OR_BC_INLET) {
                const double *ax = &c->bareav[12 * f + 3 * ip];
                double rip = 0;
                for (int ic = 0; ic < npf; ++ic) rip += fsf[npf * ip + ic] * c->rho[sd.fn[ic]];
                double m = 0;
                for (int j = 0; j < 3; ++j) m += (rip * Ubc[j]) * ax[j];
                *md = urf * m + (1.0 - urf) * *md;
            } else {
                if (t == OR_BC_OUTLET && c->rfl[4 * f + ip] == 1) {  /* slip wall while reversed */
                    *md = 0.0;
                    continue;
                }
                const double m = outlet_rhie_chow(c, f, ip, fdn, fsf, &sd, t == OR_BC_OUTLET);
                *md = urf * m + (1.0 - urf) * *md;
            }
        }
    }
}

void or_bc_reversal(or_case *c)
{
    for (int f = 0; f < c->nbf; ++f) {
        if (face_type(c, f) != OR_BC_OUTLET) continue;
        or_side sd;
        or_side_get(c, f, &sd);
        const int npf = sd.npf;
        const double fw = 1.0 / npf;
        double *md = &c->msd[4 * f];
        int *rf = &c->rfl[4 * f];
        double side = 0;
        for (int ip = 0; ip < npf; ++ip) {
            side += md[ip];
            md[ip] = fmax(md[ip], 0.0);
        }
        const double pbc = c->bc_p[c->bf_bc[f]];
        for (int ip = 0; ip < npf; ++ip) {
            const double *ax = &c->bareav[12 * f + 3 * ip];
            const double amag = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            double pavg = 0, uavg[3] = {0, 0, 0};
            for (int ic = 0; ic < npf; ++ic) {
                pavg += fw * c->p[sd.fn[ic]];
                for (int j = 0; j < 3; ++j) uavg[j] += fw * c->U[3 * sd.fn[ic] + j];
            }
            double vn = 0;
            for (int j = 0; j < 3; ++j) vn += uavg[j] * ax[j] / amag;
            if (rf[ip] == 0) {
                if (side < -SMALL) {
                    rf[ip] = 1;
                    md[ip] = 0.0;
                }
            } else if (vn >= 0.0 && pbc <= pavg) {
                rf[ip] = 0;
            }
        }
    }
}

void or_bc_divergence(or_case *c)
{
    for (int f = 0; f < c->nbf; ++f) {
        if (!flow_side(face_type(c, f))) continue;
        or_side sd;
        or_side_get(c, f, &sd);
        for (int ip = 0; ip < sd.npf; ++ip) c->div[sd.fn[ip]] += c->msd[4 * f + ip];
    }
}

/* ------------------------------------------------------------ momentum */
/* element lhs [3 npe][3 npe] (row-major), rhs [3 npe] */
void or_bc_momentum(or_case *c, int f, double *lhs, double *rhs)
{
    const or_params *P = &c->prm;
    const int t = face_type(c, f);
    or_side sd;
    double co[24], fdn[96], fdet[4], fsf[16], Uw[24], bcm[8], muf[4];
    side_gather(c, f, &sd, co);
    const int npf = sd.npf, npe = sd.npe, ld = 3 * npe;
    or_side_shape(npf, P->u_shifted, fsf);
    or_topo_side_gradop(sd.topo, P->ugrad_shifted, sd.ord, co, fdn, fdet);
    const double *Ubc = &c->bc_U[3 * c->bf_bc[f]];
    for (int a = 0; a < npe; ++a) {
        bcm[a] = 1.0;
        for (int d = 0; d < 3; ++d) Uw[3 * a + d] = c->U[3 * sd.en[a] + d];
    }
    for (int k = 0; k < npf; ++k) {
        muf[k] = c->mu[sd.fn[k]];
        if (t == OR_BC_INLET) {
            const int ic = sd.sn[k];
            bcm[ic] = 0.0;
            for (int d = 0; d < 3; ++d) Uw[3 * ic + d] = Ubc[d];
        }
    }
    for (int ip = 0; ip < npf; ++ip) {
        const int nearest = sd.sn[ip];
        const double *ax = &c->bareav[12 * f + 3 * ip];
        const double tm = c->msd[4 * f + ip];
        double mub = 0;
        for (int ic = 0; ic < npf; ++ic) mub += fsf[npf * ip + ic] * muf[ic];
        if (t == OR_BC_INLET) {
            for (int i = 0; i < 3; ++i) rhs[3 * nearest + i] -= tm * Ubc[i];
            for (int ic = 0; ic < npe; ++ic) {
                const double *dn = &fdn[3 * npe * ip + 3 * ic];
                for (int i = 0; i < 3; ++i) {
                    const int iR = 3 * nearest + i, rowR = iR * ld;
                    for (int j = 0; j < 3; ++j) {
                        double lf = -mub * dn[j] * ax[j];
                        lhs[rowR + 3 * ic + i] += lf * bcm[ic];
                        rhs[iR] -= lf * Uw[3 * ic + i];
                        lf = -mub * dn[i] * ax[j];
                        lhs[rowR + 3 * ic + j] += lf * bcm[ic];
                        rhs[iR] -= lf * Uw[3 * ic + j];
                    }
                }
            }
            continue;
        }
        if (t == OR_BC_OUTLET && c->rfl[4 * f + ip] == 1) continue;  /* reversed: slip wall */
        const double amag = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        double nx[3];
        for (int i = 0; i < 3; ++i) nx[i] = ax[i] / amag;
        /* an opening ip with msd <= 0 entrains: explicit flux msd Iu d and
         * the zero normal stress without the grad(U)^T terms */
        const int entrain = t == OR_BC_OPENING && !(tm > 0.0);
        if (entrain) {
            double d[3], uip[3] = {0, 0, 0}, num = 0, den = 0;
            opening_dir(c, f, ip, d);
            for (int ic = 0; ic < npf; ++ic)
                for (int j = 0; j < 3; ++j) uip[j] += fsf[npf * ip + ic] * c->U[3 * sd.fn[ic] + j];
            for (int i = 0; i < 3; ++i) {
                num += uip[i] * nx[i];
                den += d[i] * nx[i];
            }
            const double Iu = num / (den + SMALL);
            for (int i = 0; i < 3; ++i) rhs[3 * nearest + i] -= tm * Iu * d[i];
        } else {
            for (int i = 0; i < 3; ++i) {
                const int iR = 3 * nearest + i;
                rhs[iR] -= tm * Uw[3 * nearest + i];
                lhs[iR * ld + 3 * nearest + i] += tm;
            }
        }
        for (int ic = 0; ic < npe; ++ic) {
            const double *dn = &fdn[3 * npe * ip + 3 * ic];
            for (int j = 0; j < 3; ++j) {
                const double axj = ax[j], dnj = dn[j], uxj = Uw[3 * ic + j];
                for (int i = 0; i < 3; ++i) {
                    const int iR = 3 * nearest + i, rowR = iR * ld;
                    const double dni = dn[i], uxi = Uw[3 * ic + i], nxi = nx[i];
                    const double om = 1.0 - nxi * nxi;
                    double lf = -mub * dnj * axj * om;
                    lhs[rowR + 3 * ic + i] += lf;
                    rhs[iR] -= lf * uxi;
                    if (!entrain) {
                        lf = -mub * dni * axj * om;
                        lhs[rowR + 3 * ic + j] += lf;
                        rhs[iR] -= lf * uxj;
                    }
                    for (int l = 0; l < 3; ++l) {
                        if (l == i) continue;
                        const double nn_ = nxi * nx[l];
                        lf = mub * dnj * axj * nn_;
                        lhs[rowR + 3 * ic + l] += lf;
                        rhs[iR] -= lf * Uw[3 * ic + l];
                        if (entrain) continue;
                        lf = mub * dn[l] * axj * nn_;
                        lhs[rowR + 3 * ic + j] += lf;
                        rhs[iR] -= lf * uxj;
                    }
                }
            }
        }
    }
}

/* --------------------------------------------------- pressure correction */
/* element lhs [npe][npe], rhs [npe] */
void or_bc_pcorr(or_case *c, int f, double *lhs, double *rhs)
{
    const or_params *P = &c->prm;
    const int t = face_type(c, f);
    or_side sd;
    double co[24], fdn[96], fdet[4], fsf[16], bcm[8];
    side_gather(c, f, &sd, co);
    const int npf = sd.npf, npe = sd.npe;
    or_side_shape(npf, P->u_shifted, fsf);
    const double *Ubc = &c->bc_U[3 * c->bf_bc[f]];
    if (t != OR_BC_INLET) or_topo_side_gradop(sd.topo, P->pgrad_shifted, sd.ord, co, fdn, fdet);
    for (int a = 0; a < npe; ++a) bcm[a] = 1.0;
    for (int k = 0; k < npf; ++k) bcm[sd.sn[k]] = 0.0;
    for (int ip = 0; ip < npf; ++ip) {
        const int nearest = sd.sn[ip];
        const double *ax = &c->bareav[12 * f + 3 * ip];
        double rip = 0;
        for (int ic = 0; ic < npf; ++ic) rip += fsf[npf * ip + ic] * c->rho[sd.fn[ic]];
        if (t == OR_BC_INLET) {
            double m = 0;
            for (int j = 0; j < 3; ++j) m += (rip * Ubc[j]) * ax[j];
            rhs[nearest] -= m;
            continue;
        }
        /* reversed outlet ip: no mass flux (static mesh); openings: every ip */
        if (t == OR_BC_OUTLET && c->rfl[4 * f + ip] != 0) continue;
        double du[3] = {0, 0, 0};
        for (int ic = 0; ic < npf; ++ic)
            for (int j = 0; j < 3; ++j) du[j] += fsf[npf * ip + ic] * c->du[3 * sd.fn[ic] + j];
        for (int ic = 0; ic < npe; ++ic) {
            double lf = 0;
            for (int j = 0; j < 3; ++j) lf += -rip * du[j] * fdn[3 * npe * ip + 3 * ic + j] * ax[j];
            lhs[npe * nearest + ic] += lf * bcm[ic];
        }
        rhs[nearest] -= outlet_rhie_chow(c, f, ip, fdn, fsf, &sd, 1);
    }
}

/* -------------------------------------------------------- k and omega */
/* element lhs [npe][npe], rhs [npe]; G8: the diffusivity at the element's nodes */
void or_bc_scalar(or_case *c, int which, int f, const double *G8, double *lhs, double *rhs)
{
    const or_sst_params *S = &c->sst;
    const int t = face_type(c, f);
    const double *phiF = which == 0 ? c->k : c->w;
    or_side sd;
    double co[24], fdn[96], fdet[4], fsf[16], bcm[8], ph[8];
    side_gather(c, f, &sd, co);
    const int npf = sd.npf, npe = sd.npe;
    const double pbc = which == 0 ? c->bc_k[c->bf_bc[f]] : c->bc_w[c->bf_bc[f]];
    for (int a = 0; a < npe; ++a) {
        bcm[a] = 1.0;
        ph[a] = phiF[sd.en[a]];
    }
    for (int k = 0; k < npf; ++k) {
        bcm[sd.sn[k]] = 0.0;
        ph[sd.sn[k]] = pbc;
    }
    if (t == OR_BC_INLET) {
        or_side_shape(npf, which == 0 ? S->k_shifted : S->w_shifted, fsf);
        or_topo_side_gradop(sd.topo, which == 0 ? S->kgrad_shifted : S->wgrad_shifted, sd.ord, co,
                            fdn, fdet);
    }
    for (int ip = 0; ip < npf; ++ip) {
        if (c->rfl[4 * f + ip] != 0) continue;
        const int nearest = sd.sn[ip];
        const double tm = c->msd[4 * f + ip];
        if (t != OR_BC_INLET) {  /* outlet, opening: implicit upwind */
            rhs[nearest] -= tm * phiF[sd.en[nearest]];
            lhs[npe * nearest + nearest] += tm;
            continue;
        }
        rhs[nearest] -= tm * pbc;
        double Gb = 0;
        for (int ic = 0; ic < npf; ++ic) Gb += fsf[npf * ip + ic] * G8[sd.sn[ic]];
        const double *ax = &c->bareav[12 * f + 3 * ip];
        for (int ic = 0; ic < npe; ++ic)
            for (int j = 0; j < 3; ++j) {
                const double lf = -Gb * fdn[3 * npe * ip + 3 * ic + j] * ax[j];
                lhs[npe * nearest + ic] += lf * bcm[ic];
                rhs[nearest] -= lf * ph[ic];
            }
    }
}
