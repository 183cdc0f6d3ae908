// Sibson natural-neighbour interpolation at interior query points, on the
// host: the inner loop of MATLAB's scatteredInterpolant(..., 'natural') as
// the reference calls it after the global step
// (func_quadtree/funGlobalNodalStrainQuadtree.m:285-292) and in the refill of
// the local steps (func_quadtree/LocalICGNQuadtree.m:313-330).  The caller
// supplies the Delaunay triangulation of the sites, each site's Delaunay
// neighbours and the triangle that holds each query; queries on a site, on
// the hull boundary or outside the hull are the caller's (they take a site
// value, a linear edge value or the nearest site).
//
// For a query q: the natural-neighbour candidates are the vertices of the
// triangles whose circumcircle holds q (a walk over triangle neighbours from
// q's triangle); q's new Voronoi cell is a large box cut by the bisector
// half-planes of q and every candidate; the natural neighbours proper are the
// candidates whose bisector carries an edge of that cell; the area site i
// loses is the cell cut to i's old Voronoi cell (the half-planes of i and its
// Delaunay neighbours).  Weights are those areas over their sum.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "../../include/staq_dic.h"

namespace {

struct Pt {
    double x, y;
};

// convex polygon `in` (n vertices) cut to the half-plane ax*x + ay*y <= b
int clip(const Pt *in, int n, double ax, double ay, double b, Pt *out)
{
    if (n == 0) return 0;
    int m = 0;
    Pt p = in[n - 1];
    double sp = ax * p.x + ay * p.y - b;
    for (int i = 0; i < n; ++i) {
        const Pt c = in[i];
        const double sc = ax * c.x + ay * c.y - b;
        if (sc <= 0.0) {
            if (sp > 0.0) {
                const double t = sp / (sp - sc);
                out[m++] = {p.x + t * (c.x - p.x), p.y + t * (c.y - p.y)};
            }
            out[m++] = c;
        } else if (sp <= 0.0) {
            const double t = sp / (sp - sc);
            out[m++] = {p.x + t * (c.x - p.x), p.y + t * (c.y - p.y)};
        }
        p = c;
        sp = sc;
    }
    return m;
}

double area(const Pt *p, int n)
{
    if (n < 3) return 0.0;
    double s = 0.0;
    Pt a = p[n - 1];
    for (int i = 0; i < n; ++i) {
        s += a.x * p[i].y - a.y * p[i].x;
        a = p[i];
    }
    return 0.5 * std::fabs(s);
}

}  // namespace

extern "C" int staq_natural_interp(long n_sites, const double *sites, long n_tri, const int *tri,
                                   const int *nbr, const long *vptr, const int *vadj, long m,
                                   const double *values, long n_q, const double *queries,
                                   const int *qtri, double *out)
{
    if (n_sites < 3 || n_tri < 1 || m < 1 || n_q < 0 || !sites || !tri || !nbr || !vptr || !vadj ||
        !values || (n_q > 0 && (!queries || !qtri || !out)))
        return STAQ_EINVAL;
    for (long k = 0; k < n_q; ++k)
        if (qtri[k] < 0 || qtri[k] >= n_tri) return STAQ_EINVAL;
    // circumcentres relative to each triangle's first vertex, and the span
    std::vector<double> ccx(n_tri), ccy(n_tri), r2(n_tri);
    for (long t = 0; t < n_tri; ++t) {
        const double *a = sites + 2 * tri[3 * t], *b = sites + 2 * tri[3 * t + 1],
                     *c = sites + 2 * tri[3 * t + 2];
        const double bx = b[0] - a[0], by = b[1] - a[1], cx = c[0] - a[0], cy = c[1] - a[1];
        const double d = 2.0 * (bx * cy - by * cx);
        const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
        const double ux = (cy * b2 - by * c2) / d, uy = (bx * c2 - cx * b2) / d;
        ccx[t] = a[0] + ux;
        ccy[t] = a[1] + uy;
        r2[t] = ux * ux + uy * uy;
    }
    double xmin = sites[0], xmax = sites[0], ymin = sites[1], ymax = sites[1];
    for (long i = 1; i < n_sites; ++i) {
        xmin = std::fmin(xmin, sites[2 * i]);
        xmax = std::fmax(xmax, sites[2 * i]);
        ymin = std::fmin(ymin, sites[2 * i + 1]);
        ymax = std::fmax(ymax, sites[2 * i + 1]);
    }
    double span = std::fmax(xmax - xmin, ymax - ymin);
    if (!(span > 0)) span = 1.0;
    const double big = 1e4 * span;

    std::vector<long> seen(n_tri, -1), stamp(n_sites, -1);
    std::vector<int> stack, cand, nn;
    std::vector<Pt> cell, bufa, bufb;
    std::vector<double> ar;
    for (long k = 0; k < n_q; ++k) {
        const double qx = queries[2 * k], qy = queries[2 * k + 1];
        // cavity walk
        cand.clear();
        stack.clear();
        const int s0 = qtri[k];
        seen[s0] = k;
        stack.push_back(s0);
        for (int v = 0; v < 3; ++v) {
            const int g = tri[3 * s0 + v];
            if (stamp[g] != k) {
                stamp[g] = k;
                cand.push_back(g);
            }
        }
        while (!stack.empty()) {
            const int t = stack.back();
            stack.pop_back();
            for (int e = 0; e < 3; ++e) {
                const int u = nbr[3 * t + e];
                if (u < 0 || seen[u] == k) continue;
                seen[u] = k;
                const double dx = qx - ccx[u], dy = qy - ccy[u];
                if (dx * dx + dy * dy <= r2[u] * (1.0 + 1e-10)) {
                    stack.push_back(u);
                    for (int v = 0; v < 3; ++v) {
                        const int g = tri[3 * u + v];
                        if (stamp[g] != k) {
                            stamp[g] = k;
                            cand.push_back(g);
                        }
                    }
                }
            }
        }
        // q's cell
        const size_t cap = 8 + cand.size() + 64;
        cell.assign(cap, Pt{0, 0});
        bufa.assign(cap, Pt{0, 0});
        bufb.assign(cap, Pt{0, 0});
        int nc = 4;
        cell[0] = {qx - big, qy - big};
        cell[1] = {qx + big, qy - big};
        cell[2] = {qx + big, qy + big};
        cell[3] = {qx - big, qy + big};
        for (int j : cand) {
            const double px = sites[2 * j], py = sites[2 * j + 1];
            const double ax = px - qx, ay = py - qy;
            nc = clip(cell.data(), nc, ax, ay, ax * 0.5 * (px + qx) + ay * 0.5 * (py + qy), bufa.data());
            std::copy(bufa.begin(), bufa.begin() + nc, cell.begin());
        }
        // natural neighbours proper: two cell vertices on the bisector
        nn.clear();
        for (int j : cand) {
            const double px = sites[2 * j], py = sites[2 * j + 1];
            const double ax = px - qx, ay = py - qy;
            const double b = 0.5 * ((px + qx) * ax + (py + qy) * ay);
            const double tol = 1e-9 * span * std::sqrt(ax * ax + ay * ay);
            int on = 0;
            for (int v = 0; v < nc; ++v)
                if (std::fabs(cell[v].x * ax + cell[v].y * ay - b) <= tol) ++on;
            if (on >= 2) nn.push_back(j);
        }
        // the area each natural neighbour loses
        ar.assign(nn.size(), 0.0);
        double tot = 0.0;
        for (size_t i = 0; i < nn.size(); ++i) {
            const int gi = nn[i];
            const double xi = sites[2 * gi], yi = sites[2 * gi + 1];
            const long deg = vptr[gi + 1] - vptr[gi];
            const size_t need = (size_t)nc + (size_t)deg + 8;
            if (bufa.size() < need) {
                bufa.resize(need);
                bufb.resize(need);
            }
            int np = nc;
            std::copy(cell.begin(), cell.begin() + nc, bufa.begin());
            for (long e = vptr[gi]; e < vptr[gi + 1] && np >= 3; ++e) {
                const int gj = vadj[e];
                const double xj = sites[2 * gj], yj = sites[2 * gj + 1];
                const double ax = xj - xi, ay = yj - yi;
                np = clip(bufa.data(), np, ax, ay, ax * 0.5 * (xj + xi) + ay * 0.5 * (yj + yi),
                          bufb.data());
                std::swap(bufa, bufb);
            }
            ar[i] = area(bufa.data(), np);
            tot += ar[i];
        }
        for (long c = 0; c < m; ++c) {
            double s = 0.0;
            for (size_t i = 0; i < nn.size(); ++i)
                if (ar[i] > 0) s += (ar[i] / tot) * values[nn[i] * m + c];
            out[k * m + c] = s;
        }
    }
    return STAQ_OK;
}
