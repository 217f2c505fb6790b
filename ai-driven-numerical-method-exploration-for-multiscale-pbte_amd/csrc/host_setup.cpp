// host_setup.cpp -- host-side set-up of a case: the gmsh 2.2 ASCII reader
// and the face integrals of the nodal DG basis (the element x face-basis
// and element x neighbour tensors), threaded over elements.
//
// The face integrals repeat, element by element, the arithmetic of the
// vectorised construction they replace (solver-python/pybte/integration.py:
// 321-404): the same expressions in the same order, so every value is the
// same double.  One exception: a monomial power x^e with e >= 3 is libm's
// pow(x, e), which numpy's power also calls unless the CPU has AVX-512, in
// which case numpy uses its own vector pow and a value may differ by one
// unit in the last place; e <= 2 is 1, x or x * x, as numpy's is.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "pbte_hip.h"

#pragma clang fp contract(off)

namespace {

int host_threads(int asked)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int cap = (int)std::min(16u, hw);
    return asked > 0 ? std::min(asked, cap) : cap;
}

template <typename F> void parallel_range(int64_t n, int threads, F&& fn)
{
    if (threads <= 1 || n < 2048) {
        fn(0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    const int64_t chunk = 1024;
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&]() {
            for (int64_t s = next.fetch_add(chunk); s < n; s = next.fetch_add(chunk)) fn(s, std::min(n, s + chunk));
        });
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// gmsh 2.2 ASCII
// ---------------------------------------------------------------------------
struct Cursor {
    const char* p;
    const char* end;
    bool eol() const { return p >= end || *p == '\n'; }
    void skip_blank()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    }
    void next_line()
    {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    // the rest of the line, stripped
    std::string line()
    {
        skip_blank();
        const char* s = p;
        while (p < end && *p != '\n') ++p;
        const char* e = p;
        while (e > s && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\r')) --e;
        if (p < end) ++p;
        return std::string(s, e);
    }
    bool integer(long long& v)
    {
        skip_blank();
        if (eol()) return false;
        char* q = nullptr;
        errno = 0;
        v = std::strtoll(p, &q, 10);
        if (q == p || errno) return false;
        if (q < end && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n') return false;
        p = q;
        return true;
    }
    bool real(double& v)
    {
        skip_blank();
        if (eol()) return false;
        char* q = nullptr;
        v = std::strtod(p, &q);
        if (q == p) return false;
        p = q;
        return true;
    }
};

}  // namespace

struct pbte_gmsh {
    std::vector<double> xy;              // [n_nodes][2]
    std::vector<int32_t> type, phyid;    // per element
    std::vector<int32_t> conn;           // [n_elem][3], -1 after a line's two nodes
    std::string names;                   // the $PhysicalNames item lines, one per line
};

extern "C" pbte_gmsh* pbte_gmsh_read(const char* path, char* err, int32_t errlen)
{
    auto fail = [&](const std::string& m) -> pbte_gmsh* {
        if (err && errlen > 0) {
            std::strncpy(err, m.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    };
    std::string text;
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) return fail(std::string(path) + ": cannot open");
        std::ostringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    }
    const std::string P = path;
    // section bodies: the first occurrence of each name
    struct Sec { const char* b = nullptr; const char* e = nullptr; };
    Sec fmt, names, nodes, elems;
    {
        Cursor c{text.data(), text.data() + text.size()};
        std::string open;
        const char* body = nullptr;
        while (c.p < c.end) {
            const char* ls = c.p;
            const std::string line = c.line();
            if (!open.empty()) {
                if (line == "$End" + open) {
                    Sec* s = open == "MeshFormat" ? &fmt : open == "PhysicalNames" ? &names
                           : open == "Nodes" ? &nodes : open == "Elements" ? &elems : nullptr;
                    if (s && !s->b) *s = Sec{body, ls};
                    open.clear();
                }
            } else if (!line.empty() && line[0] == '$' && line.compare(0, 4, "$End") != 0) {
                open = line.substr(1);
                body = c.p;
            }
        }
    }
    pbte_gmsh* g = new pbte_gmsh();
    try {
        if (fmt.b) {
            Cursor c{fmt.b, fmt.e};
            std::istringstream ls(c.line());
            std::string version, ftype;
            ls >> version >> ftype;
            if (version.compare(0, 3, "2.2") != 0) throw P + ": gmsh format " + version + ", this reader takes 2.2";
            if (ftype != "0") throw P + ": binary gmsh file; this reader takes ASCII";
        }
        auto count = [&](Cursor& c, const char* name) -> long long {
            long long n = 0;
            if (c.p >= c.end) throw P + ": $" + name + " section is empty";
            if (!c.integer(n) || n < 0) throw P + ": $" + name + " count is not a number";
            c.next_line();
            return n;
        };
        auto short_of = [&](const char* name, long long want, long long got) {
            return P + ": $" + name + " announces " + std::to_string(want) + " entries and holds " +
                   std::to_string(got);
        };
        if (names.b) {
            Cursor c{names.b, names.e};
            const long long n = count(c, "PhysicalNames");
            for (long long k = 0; k < n; ++k) {
                if (c.p >= c.end) throw short_of("PhysicalNames", n, k);
                g->names += c.line();
                g->names += '\n';
            }
        }
        if (!nodes.b) throw P + ": no $Nodes section";
        {
            Cursor c{nodes.b, nodes.e};
            const long long n = count(c, "Nodes");
            g->xy.resize((size_t)n * 2);
            for (long long k = 0; k < n; ++k) {
                if (c.p >= c.end) throw short_of("Nodes", n, k);
                long long id = 0;
                double x = 0, y = 0, z = 0;
                if (!c.integer(id) || !c.real(x) || !c.real(y) || !c.real(z))
                    throw P + ": node line " + std::to_string(k + 1) + " does not hold an id and three coordinates";
                if (id != k + 1) throw P + ": node ids do not run 1.." + std::to_string(n) + " in file order";
                g->xy[(size_t)k * 2] = x;
                g->xy[(size_t)k * 2 + 1] = y;
                c.next_line();
            }
        }
        if (elems.b) {
            Cursor c{elems.b, elems.e};
            const long long n = count(c, "Elements");
            g->type.resize((size_t)n);
            g->phyid.resize((size_t)n);
            g->conn.assign((size_t)n * 3, -1);
            for (long long k = 0; k < n; ++k) {
                if (c.p >= c.end) throw short_of("Elements", n, k);
                long long eid = 0, kind = 0, ntags = 0;
                if (!c.integer(eid) || !c.integer(kind) || !c.integer(ntags))
                    throw P + ": element line " + std::to_string(k + 1) + " is not a gmsh element";
                const int width = kind == 1 ? 2 : kind == 2 ? 3 : 0;
                if (!width)
                    throw P + ": element " + std::to_string(eid) + " is of type " + std::to_string(kind) +
                        "; only lines (1) and triangles (2) are read";
                if (ntags < 2)
                    throw P + ": element " + std::to_string(eid) + " carries " + std::to_string(ntags) +
                        " tags; the physical and geometrical ones are both needed";
                long long v = 0;
                for (long long t = 0; t < ntags; ++t) {
                    if (!c.integer(v)) throw P + ": element " + std::to_string(eid) + " is short of tags";
                    if (t == 0) g->phyid[(size_t)k] = (int32_t)v;
                }
                for (int w = 0; w < width; ++w) {
                    if (!c.integer(v)) throw P + ": element " + std::to_string(eid) + " is short of nodes";
                    g->conn[(size_t)k * 3 + w] = (int32_t)(v - 1);
                }
                g->type[(size_t)k] = (int32_t)kind;
                c.next_line();
            }
        }
    } catch (const std::string& m) {
        delete g;
        return fail(m);
    }
    return g;
}

extern "C" void pbte_gmsh_free(pbte_gmsh* g) { delete g; }

extern "C" int pbte_gmsh_sizes(const pbte_gmsh* g, int64_t* out3)
{
    if (!g || !out3) return -1;
    out3[0] = (int64_t)g->xy.size() / 2;
    out3[1] = (int64_t)g->type.size();
    out3[2] = (int64_t)g->names.size();
    return 0;
}

extern "C" int pbte_gmsh_copy(const pbte_gmsh* g, double* xy, int32_t* type, int32_t* phyid, int32_t* conn,
                              char* names)
{
    if (!g) return -1;
    if (xy) std::copy(g->xy.begin(), g->xy.end(), xy);
    if (type) std::copy(g->type.begin(), g->type.end(), type);
    if (phyid) std::copy(g->phyid.begin(), g->phyid.end(), phyid);
    if (conn) std::copy(g->conn.begin(), g->conn.end(), conn);
    if (names) std::memcpy(names, g->names.data(), g->names.size());
    return 0;
}

// ---------------------------------------------------------------------------
// face integrals
// ---------------------------------------------------------------------------
namespace {

inline double mono_pow(double x, int e)
{
    if (e == 0) return 1.0;
    if (e == 1) return x;
    if (e == 2) return x * x;
    return std::pow(x, (double)e);
}

// physical -> reference map of a triangle: xi = b x + c y + d, eta = f x + g y + h
inline void affine(const double* x, const double* y, double* m)
{
    const double x1 = x[0], x2 = x[1], x3 = x[2], y1 = y[0], y2 = y[1], y3 = y[2];
    const double A = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    const double B = y3 - y1;
    const double C = x1 - x3;
    const double D = (x3 - x1) * y1 - (y3 - y1) * x1;
    const double E = (x3 - x1) * (y2 - y1) - (x2 - x1) * (y3 - y1);
    const double F = y2 - y1;
    const double G = x1 - x2;
    const double H = (x2 - x1) * y1 - (y2 - y1) * x1;
    m[0] = B / A; m[1] = C / A; m[2] = D / A;
    m[3] = F / E; m[4] = G / E; m[5] = H / E;
}

// the P basis values at (xi, eta): sum over the monomials k in order of
// (C[m][k] xi^a_k) eta^b_k
inline void eval_mono(const pbte_face_integrals_desc* d, double xi, double eta, double* out)
{
    const int P = d->nd;
    for (int m = 0; m < P; ++m) out[m] = 0.0;
    for (int k = 0; k < P; ++k) {
        const double pa = mono_pow(xi, d->mono[2 * k]), pb = mono_pow(eta, d->mono[2 * k + 1]);
        for (int m = 0; m < P; ++m) out[m] += (d->C[m * P + k] * pa) * pb;
    }
}

}  // namespace

extern "C" int pbte_face_integrals(const pbte_face_integrals_desc* d, double* tri_fc_fc, double* tri_tri_fc,
                                   int32_t threads)
{
    if (!d || !tri_fc_fc || !tri_tri_fc) return -1;
    const int P = d->nd, Q = d->nf, NQ = d->np_fc;
    const int64_t N = d->n_tris;
    if (P < 1 || P > 15 || Q < 1 || Q > 5 || NQ < 1 || NQ > 64 || N < 0) return -1;
    const double n1 = (double)NQ, n2 = (double)(NQ + 1);
    const double n22 = (double)((NQ + 1) * (NQ + 1));
    (void)n2;
    parallel_range(N * 3, host_threads(threads), [&](int64_t g0, int64_t g1) {
        std::vector<double> pm((size_t)NQ * P), pn((size_t)NQ * P), pl((size_t)NQ * Q), w(NQ), ab(NQ);
        double xs[3], ys[3], xo[3], yo[3], ma[6], mb[6];
        for (int64_t g = g0; g < g1; ++g) {
            const int64_t i = g / 3;
            const int j2 = (int)(g - i * 3);
            const int f = d->tri_faces[i * 3 + j2];
            const double ln = d->face_len[f];
            for (int q = 0; q < NQ; ++q) {
                const double yq = d->yq[q], lp = d->lpq[q];
                ab[q] = (0.0 * (1.0 - yq) + ln * (1.0 + yq)) / 2.0;
                w[q] = n22 * (ln - 0.0) / ((1.0 - yq * yq) * lp * lp) / n1 / n1;
            }
            const int a0 = d->face_nodes[2 * f], a1 = d->face_nodes[2 * f + 1];
            const double fx1 = d->nodes[2 * a0], fy1 = d->nodes[2 * a0 + 1];
            const double fx2 = d->nodes[2 * a1], fy2 = d->nodes[2 * a1 + 1];
            for (int k = 0; k < 3; ++k) {
                const int nd = d->tri_nodes[i * 3 + k];
                xs[k] = d->nodes[2 * nd];
                ys[k] = d->nodes[2 * nd + 1];
            }
            affine(xs, ys, ma);
            const int o = d->other[i * 3 + j2];
            if (o >= 0) {
                for (int k = 0; k < 3; ++k) {
                    const int nd = d->tri_nodes[(int64_t)o * 3 + k];
                    xo[k] = d->nodes[2 * nd] + d->off_x[i * 3 + j2];
                    yo[k] = d->nodes[2 * nd + 1] + d->off_y[i * 3 + j2];
                }
                affine(xo, yo, mb);
            }
            for (int q = 0; q < NQ; ++q) {
                const double xp = (fx2 - fx1) / ln * ab[q] + fx1;
                const double yp = (fy2 - fy1) / ln * ab[q] + fy1;
                eval_mono(d, ma[0] * xp + ma[1] * yp + ma[2], ma[3] * xp + ma[4] * yp + ma[5], &pm[(size_t)q * P]);
                if (o >= 0)
                    eval_mono(d, mb[0] * xp + mb[1] * yp + mb[2], mb[3] * xp + mb[4] * yp + mb[5],
                              &pn[(size_t)q * P]);
                const double t = 2.0 * ab[q] / ln - 1.0;
                double* plq = &pl[(size_t)q * Q];
                for (int l = 0; l < Q; ++l) plq[l] = 0.0;
                for (int k = 0; k < Q; ++k) {
                    const double tk = mono_pow(t, k);
                    for (int l = 0; l < Q; ++l) plq[l] += d->Cf[l * Q + k] * tk;
                }
            }
            // element x face basis: int_tri_fc_fc[m][l][i][j2]
            for (int m = 0; m < P; ++m)
                for (int l = 0; l < Q; ++l) {
                    double acc = 0.0;
                    for (int q = 0; q < NQ; ++q) acc += (pm[(size_t)q * P + m] * pl[(size_t)q * Q + l]) * w[q];
                    tri_fc_fc[(((size_t)m * Q + l) * N + i) * 3 + j2] = acc;
                }
            // element x neighbour (int_tri_tri_fc[i][j2][m][l]), or on a wall
            // int phi_m ds in every column
            double* out = tri_tri_fc + ((size_t)i * 3 + j2) * P * P;
            if (o >= 0) {
                for (int m = 0; m < P; ++m)
                    for (int l = 0; l < P; ++l) {
                        double acc = 0.0;
                        for (int q = 0; q < NQ; ++q) acc += (pm[(size_t)q * P + m] * pn[(size_t)q * P + l]) * w[q];
                        out[m * P + l] = acc;
                    }
            } else {
                for (int m = 0; m < P; ++m) {
                    double acc = 0.0;
                    for (int q = 0; q < NQ; ++q) acc += pm[(size_t)q * P + m] * w[q];
                    for (int l = 0; l < P; ++l) out[m * P + l] = acc;
                }
            }
        }
    });
    return 0;
}
