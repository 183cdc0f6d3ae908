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

// Per-query Sibson coordinates over one triangulation: the circumcircles are
// computed once, then query() walks, clips and measures.
class Sibson {
public:
    Sibson(long n_sites, const double *sites, long n_tri, const int *tri, const int *nbr,
           const long *vptr, const int *vadj)
        : sites_(sites), tri_(tri), nbr_(nbr), vptr_(vptr), vadj_(vadj), ccx_(n_tri), ccy_(n_tri),
          r2_(n_tri), seen_(n_tri, -1), stamp_(n_sites, -1)
    {
        // circumcentres relative to each triangle's first vertex, and the span
        for (long t = 0; t < n_tri; ++t) {
            const double *a = sites + 2 * tri[3 * t], *b = sites + 2 * tri[3 * t + 1],
                         *c = sites + 2 * tri[3 * t + 2];
            const double bx = b[0] - a[0], by = b[1] - a[1], cx = c[0] - a[0], cy = c[1] - a[1];
            const double d = 2.0 * (bx * cy - by * cx);
            const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
            const double ux = (cy * b2 - by * c2) / d, uy = (bx * c2 - cx * b2) / d;
            ccx_[t] = a[0] + ux;
            ccy_[t] = a[1] + uy;
            r2_[t] = ux * ux + uy * uy;
        }
        double xmin = sites[0], xmax = sites[0], ymin = sites[1], ymax = sites[1];
        for (long i = 1; i < n_sites; ++i) {
            xmin = std::fmin(xmin, sites[2 * i]);
            xmax = std::fmax(xmax, sites[2 * i]);
            ymin = std::fmin(ymin, sites[2 * i + 1]);
            ymax = std::fmax(ymax, sites[2 * i + 1]);
        }
        span_ = std::fmax(xmax - xmin, ymax - ymin);
        if (!(span_ > 0)) span_ = 1.0;
    }

    // natural neighbours of query k at (qx, qy) in triangle s0: their site
    // indices in nn and the areas they lose in ar (zero areas included)
    void query(long k, double qx, double qy, int s0, std::vector<int> &nn, std::vector<double> &ar)
    {
        const double big = 1e4 * span_;
        const double *sites = sites_;
        // cavity walk
        cand_.clear();
        stack_.clear();
        seen_[s0] = k;
        stack_.push_back(s0);
        for (int v = 0; v < 3; ++v) {
            const int g = tri_[3 * s0 + v];
            if (stamp_[g] != k) {
                stamp_[g] = k;
                cand_.push_back(g);
            }
        }
        while (!stack_.empty()) {
            const int t = stack_.back();
            stack_.pop_back();
            for (int e = 0; e < 3; ++e) {
                const int u = nbr_[3 * t + e];
                if (u < 0 || seen_[u] == k) continue;
                seen_[u] = k;
                const double dx = qx - ccx_[u], dy = qy - ccy_[u];
                if (dx * dx + dy * dy <= r2_[u] * (1.0 + 1e-10)) {
                    stack_.push_back(u);
                    for (int v = 0; v < 3; ++v) {
                        const int g = tri_[3 * u + v];
                        if (stamp_[g] != k) {
                            stamp_[g] = k;
                            cand_.push_back(g);
                        }
                    }
                }
            }
        }
        // q's cell
        const size_t cap = 8 + cand_.size() + 64;
        cell_.assign(cap, Pt{0, 0});
        bufa_.assign(cap, Pt{0, 0});
        bufb_.assign(cap, Pt{0, 0});
        int nc = 4;
        cell_[0] = {qx - big, qy - big};
        cell_[1] = {qx + big, qy - big};
        cell_[2] = {qx + big, qy + big};
        cell_[3] = {qx - big, qy + big};
        for (int j : cand_) {
            const double px = sites[2 * j], py = sites[2 * j + 1];
            const double ax = px - qx, ay = py - qy;
            nc = clip(cell_.data(), nc, ax, ay, ax * 0.5 * (px + qx) + ay * 0.5 * (py + qy), bufa_.data());
            std::copy(bufa_.begin(), bufa_.begin() + nc, cell_.begin());
        }
        // natural neighbours proper: two cell vertices on the bisector
        nn.clear();
        for (int j : cand_) {
            const double px = sites[2 * j], py = sites[2 * j + 1];
            const double ax = px - qx, ay = py - qy;
            const double b = 0.5 * ((px + qx) * ax + (py + qy) * ay);
            const double tol = 1e-9 * span_ * std::sqrt(ax * ax + ay * ay);
            int on = 0;
            for (int v = 0; v < nc; ++v)
                if (std::fabs(cell_[v].x * ax + cell_[v].y * ay - b) <= tol) ++on;
            if (on >= 2) nn.push_back(j);
        }
        // the area each natural neighbour loses
        ar.assign(nn.size(), 0.0);
        for (size_t i = 0; i < nn.size(); ++i) {
            const int gi = nn[i];
            const double xi = sites[2 * gi], yi = sites[2 * gi + 1];
            const long deg = vptr_[gi + 1] - vptr_[gi];
            const size_t need = (size_t)nc + (size_t)deg + 8;
            if (bufa_.size() < need) {
                bufa_.resize(need);
                bufb_.resize(need);
            }
            int np = nc;
            std::copy(cell_.begin(), cell_.begin() + nc, bufa_.begin());
            for (long e = vptr_[gi]; e < vptr_[gi + 1] && np >= 3; ++e) {
                const int gj = vadj_[e];
                const double xj = sites[2 * gj], yj = sites[2 * gj + 1];
                const double ax = xj - xi, ay = yj - yi;
                np = clip(bufa_.data(), np, ax, ay, ax * 0.5 * (xj + xi) + ay * 0.5 * (yj + yi),
                          bufb_.data());
                std::swap(bufa_, bufb_);
            }
            ar[i] = area(bufa_.data(), np);
        }
    }

private:
    const double *sites_;
    const int *tri_, *nbr_;
    const long *vptr_;
    const int *vadj_;
    std::vector<double> ccx_, ccy_, r2_;
    std::vector<long> seen_, stamp_;
    std::vector<int> stack_, cand_;
    std::vector<Pt> cell_, bufa_, bufb_;
    double span_ = 1.0;
};

bool valid_tri(long n_sites, const double *sites, long n_tri, const int *tri, const int *nbr,
               const long *vptr, const int *vadj, long n_q, const double *queries, const int *qtri)
{
    if (n_sites < 3 || n_tri < 1 || n_q < 0 || !sites || !tri || !nbr || !vptr || !vadj ||
        (n_q > 0 && (!queries || !qtri)))
        return false;
    for (long k = 0; k < n_q; ++k)
        if (qtri[k] < 0 || qtri[k] >= n_tri) return false;
    return true;
}

}  // namespace

extern "C" int staq_natural_interp(long n_sites, const double *sites, long n_tri, const int *tri,
                                   const int *nbr, const long *vptr, const int *vadj, long m,
                                   const double *values, long n_q, const double *queries,
                                   const int *qtri, double *out)
{
    if (!valid_tri(n_sites, sites, n_tri, tri, nbr, vptr, vadj, n_q, queries, qtri) || m < 1 ||
        !values || (n_q > 0 && !out))
        return STAQ_EINVAL;
    Sibson S(n_sites, sites, n_tri, tri, nbr, vptr, vadj);
    std::vector<int> nn;
    std::vector<double> ar;
    for (long k = 0; k < n_q; ++k) {
        S.query(k, queries[2 * k], queries[2 * k + 1], qtri[k], nn, ar);
        double tot = 0.0;
        for (double a : ar) tot += a;
        for (long c = 0; c < m; ++c) {
            double s = 0.0;
            for (size_t i = 0; i < nn.size(); ++i)
                if (ar[i] > 0) s += (ar[i] / tot) * values[nn[i] * m + c];
            out[k * m + c] = s;
        }
    }
    return STAQ_OK;
}

extern "C" int staq_natural_weights(long n_sites, const double *sites, long n_tri, const int *tri,
                                    const int *nbr, const long *vptr, const int *vadj, long n_q,
                                    const double *queries, const int *qtri, long cap, int *idx,
                                    double *w, int *count)
{
    if (!valid_tri(n_sites, sites, n_tri, tri, nbr, vptr, vadj, n_q, queries, qtri) || cap < 1 ||
        (n_q > 0 && (!idx || !w || !count)))
        return STAQ_EINVAL;
    Sibson S(n_sites, sites, n_tri, tri, nbr, vptr, vadj);
    std::vector<int> nn;
    std::vector<double> ar;
    int rc = STAQ_OK;
    for (long k = 0; k < n_q; ++k) {
        S.query(k, queries[2 * k], queries[2 * k + 1], qtri[k], nn, ar);
        double tot = 0.0;
        for (double a : ar) tot += a;
        long c = 0;
        for (size_t i = 0; i < nn.size(); ++i) {
            if (!(ar[i] > 0)) continue;
            if (c < cap) {
                idx[k * cap + c] = nn[i];
                w[k * cap + c] = ar[i] / tot;
            }
            ++c;
        }
        count[k] = (int)c;
        if (c > cap) rc = STAQ_EINVAL;
    }
    return rc;
}
