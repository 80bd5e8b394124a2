/*
 * elem.c -- master-element tables by topology for the oracle's element
 * passes: hex8 through the hex8 routines, tet4 / wedge6 / pyramid5 through
 * topo.c, plus the SCS shape functions of the three.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Shape functions (external/nalu/src/master_element/):
 *   tet4     TetSCS::tet_shape_fcn        Tet4CVFEM.C:300-315 (shape_fcn :820,
 *            shifted_shape_fcn :829: the same function at the shifted ips)
 *   wedge6   WedSCS::wedge_shape_fcn      Wed6CVFEM.C:1151-1170 (:949, :958)
 *   pyramid5 PyrSCS::pyr_shape_fcn        Pyr5CVFEM.C:1365-1393 (standard ips),
 *            PyrSCS::shifted_pyr_shape_fcn :1396-1420 (shifted ips)
 */
#include "case_internal.h"
#include "oracle.h"

#include <float.h>
#include <math.h>

void or_topo_shape(int topo, int shifted, int npts, const double *par, double *sf)
{
    const int npe = or_topo_npe(topo);
    for (int j = 0; j < npts; ++j) {
        const double r = par[3 * j], s = par[3 * j + 1], z = par[3 * j + 2];
        double *N = &sf[npe * j];
        if (topo == OR_TET4) {
            N[0] = 1.0 - r - s - z;
            N[1] = r;
            N[2] = s;
            N[3] = z;
        } else if (topo == OR_WED6) {
            const double t = 1.0 - r - s;
            N[0] = 0.5 * t * (1.0 - z);
            N[1] = 0.5 * r * (1.0 - z);
            N[2] = 0.5 * s * (1.0 - z);
            N[3] = 0.5 * t * (1.0 + z);
            N[4] = 0.5 * r * (1.0 + z);
            N[5] = 0.5 * s * (1.0 + z);
        } else if (!shifted) {
            /* rational pyramid functions; the apex guarded as in the reference */
            const double omt = 1.0 - z;
            const double t = fabs(omt) > DBL_EPSILON ? z : 1.0 + copysign(DBL_EPSILON, omt);
            const double q = 0.25 / (1.0 - t);
            N[0] = (1.0 - r - t) * (1.0 - s - t) * q;
            N[1] = (1.0 + r - t) * (1.0 - s - t) * q;
            N[2] = (1.0 + r - t) * (1.0 + s - t) * q;
            N[3] = (1.0 - r - t) * (1.0 + s - t) * q;
            N[4] = t;
        } else {
            N[0] = 0.25 * (1.0 - r) * (1.0 - s) * (1.0 - z);
            N[1] = 0.25 * (1.0 + r) * (1.0 - s) * (1.0 - z);
            N[2] = 0.25 * (1.0 + r) * (1.0 + s) * (1.0 - z);
            N[3] = 0.25 * (1.0 - r) * (1.0 + s) * (1.0 - z);
            N[4] = z;
        }
    }
}

void or_tables_init(or_tables tab[4])
{
    for (int t = 0; t < 4; ++t) {
        tab[t].npe = or_topo_npe(t);
        tab[t].nscs = or_topo_nscs(t);
        tab[t].lr = t == OR_HEX8 ? or_hex_lrscv() : or_topo_lrscv(t);
        for (int sh = 0; sh < 2; ++sh) {
            if (t == OR_HEX8) or_hex_scs_shape(sh, tab[t].sf[sh]);
            else or_topo_shape(t, sh, tab[t].nscs, or_topo_scs_ips(t, sh), tab[t].sf[sh]);
        }
    }
}

void or_elem_areav(int topo, const double *co, double *av)
{
    if (topo == OR_HEX8) or_hex_scs_areav(co, av);
    else or_topo_scs_areav(topo, co, av);
}

int or_elem_gradop(int topo, int shifted, const double *co, double *g, double *detj)
{
    if (topo == OR_HEX8) return or_hex_scs_gradop(shifted, co, g, detj);
    return or_topo_scs_gradop(topo, shifted, co, g, detj);
}

void or_elem_scv(int topo, const double *co, double *v)
{
    if (topo == OR_HEX8) or_hex_scv_vol(co, v);
    else or_topo_scv_vol(topo, co, v);
}
