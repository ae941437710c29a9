/*
 * mesh.cpp -- host-side mesh input and preprocessing for the LLG engine.
 *
 * Reads gmsh files (MSH 2.2 and 4.1, ASCII and binary) without the gmsh library, builds the structured synthetic meshes the
 * benchmarks use, and turns a raw mesh into the preprocessed form that Mesh::mesh's constructor
 * produces in FeeLLGood (src/mesh.h:38-135): coordinates scaled, tetrahedra positively oriented
 * (src/tetra.cpp:410-424), every boundary/interface triangle present with a signed Ms jump
 * (src/mesh.cpp:121-245), nodes sorted along the longest axis (src/mesh.cpp:334-367).
 */
#include "../../include/fgmi_host.h"
#include "expr.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {
thread_local std::string g_err;

/* Face record used to find boundary and interface triangles. Node triple sorted ascending;
 * `flipped` is the parity of that sort, so two records with equal nodes and equal parity have
 * the same orientation. */
struct Face {
    std::array<int, 3> n;
    bool flipped;
    int region;
    bool surface;
    bool operator<(const Face &o) const { return n < o.n; }
};

Face make_face(int a, int b, int c, int region, bool surface)
{
    Face f;
    f.n = {a, b, c};
    int inversions = (a > b) + (b > c) + (a > c);
    std::sort(f.n.begin(), f.n.end());
    f.flipped = inversions % 2;
    f.region = region;
    f.surface = surface;
    return f;
}
} // namespace

struct fgh_mesh {
    /* raw content, file order, 0-based node indices */
    std::vector<double> pos;
    std::vector<int32_t> tets, tet_phys;
    std::vector<int32_t> tris, tri_phys;
    std::map<std::pair<int, int>, std::string> phys_names; /* (dim, tag) -> name */
    /* prepared content */
    bool prepared = false;
    std::vector<double> ppos;
    std::vector<int32_t> ptets, ptet_reg, ptris, ptri_reg;
    std::vector<double> pdMs;
    std::vector<fg_volume_region> vol;
    std::vector<fg_surface_region> surf;
    std::vector<int32_t> node_index;
    double l[3] = {0, 0, 0}, c[3] = {0, 0, 0};
};

namespace {

/* Values of a gmsh file section: whitespace-separated text (ASCII files) or native little-endian
 * binary (binary files: int 4 B, size_t 8 B, double 8 B, as gmsh writes them). */
struct Src {
    std::istream &in;
    bool bin;
    template <class T> T raw()
    {
        T v{};
        in.read(reinterpret_cast<char *>(&v), sizeof v);
        return v;
    }
    long long i32() /* int */
    {
        if (bin) return raw<int32_t>();
        long long v = 0;
        in >> v;
        return v;
    }
    long long u64() /* size_t (MSH 4.1 binary counts and tags) */
    {
        if (bin) return (long long)raw<uint64_t>();
        long long v = 0;
        in >> v;
        return v;
    }
    double f64()
    {
        if (bin) return raw<double>();
        double v = 0;
        in >> v;
        return v;
    }
    /* binary payloads start right after the newline that ends the section's text line */
    void payload()
    {
        if (bin) in.get();
    }
};

/* In a binary file, skip a section this reader does not use: up to its $End line. */
void skip_section(std::istream &in, const std::string &name)
{
    const std::string end = "$End" + name.substr(1);
    size_t k = 0;
    char ch;
    while (k < end.size() && in.get(ch)) k = ch == end[k] ? k + 1 : (ch == end[0] ? 1 : 0);
}

void read_physical_names(std::istream &in, fgh_mesh *m)
{
    int n;
    in >> n;
    for (int i = 0; i < n; i++) {
        int dim, tag;
        std::string name;
        in >> dim >> tag;
        std::getline(in, name);
        auto a = name.find('"'), b = name.rfind('"');
        if (a != std::string::npos && b > a) name = name.substr(a + 1, b - a - 1);
        m->phys_names[{dim, tag}] = name;
    }
}

int element_nodes(int type) { return type == 4 ? 4 : type == 2 ? 3 : type == 1 ? 2 : type == 15 ? 1 : -1; }

bool read_msh22(std::istream &in, fgh_mesh *m, bool bin)
{
    Src src{in, bin};
    std::string tok;
    while (in >> tok) {
        if (tok == "$PhysicalNames") {
            read_physical_names(in, m);
        } else if (tok == "$Nodes") {
            long n;
            in >> n;  // the count line is text in both forms
            src.payload();
            m->pos.assign(3 * n, 0.0);
            for (long i = 0; i < n; i++) {
                const long long tag = src.i32();
                const double x = src.f64(), y = src.f64(), z = src.f64();
                if (in.fail()) break;  // reported below
                if (tag < 1 || tag > n) {
                    g_err = "node tag out of range: " + std::to_string(tag);
                    return false;
                }
                m->pos[3 * (tag - 1)] = x;
                m->pos[3 * (tag - 1) + 1] = y;
                m->pos[3 * (tag - 1) + 2] = z;
            }
        } else if (tok == "$Elements") {
            long n;
            in >> n;
            src.payload();
            auto add = [&](int type, const std::vector<int> &tags, const int *nodes) {
                if (tags.empty() || tags[0] == 0) return; /* not in a physical group */
                if (type == 4) {
                    for (int k = 0; k < 4; k++) m->tets.push_back(nodes[k] - 1);
                    m->tet_phys.push_back(tags[0]);
                } else if (type == 2) {
                    for (int k = 0; k < 3; k++) m->tris.push_back(nodes[k] - 1);
                    m->tri_phys.push_back(tags[0]);
                }
            };
            long done = 0;
            while (done < n) {
                // text: one element per line (id type ntags tags nodes); binary: blocks of
                // elements of one type (type, count, ntags header, then id tags nodes each)
                long id = 0, follow = 1;
                int type, ntags;
                if (bin) {
                    type = (int)src.i32();
                    follow = (long)src.i32();
                    ntags = (int)src.i32();
                } else {
                    in >> id >> type >> ntags;
                }
                const int nn = element_nodes(type);
                if (nn < 0 || ntags < 0 || follow < 1 || done + follow > n) {
                    g_err = nn < 0 ? "unsupported element type " + std::to_string(type) : "malformed $Elements section";
                    return false;
                }
                for (long e = 0; e < follow; e++) {
                    if (bin) id = (long)src.i32();
                    std::vector<int> tags(ntags);
                    for (int k = 0; k < ntags; k++) tags[k] = (int)src.i32();
                    int nodes[4];
                    for (int k = 0; k < nn; k++) nodes[k] = (int)src.i32();
                    add(type, tags, nodes);
                }
                done += follow;
            }
        } else if (bin && tok.size() > 1 && tok[0] == '$' && tok.rfind("$End", 0) != 0) {
            skip_section(in, tok);
        }
        if (in.fail()) {
            g_err = "truncated or malformed gmsh file (" + tok + ")";
            return false;
        }
    }
    return true;
}

bool read_msh41(std::istream &in, fgh_mesh *m, bool bin)
{
    Src src{in, bin};
    std::string tok;
    std::map<std::pair<int, int>, std::vector<int>> ent_phys; /* (dim, entity) -> physical tags */
    while (in >> tok) {
        if (tok == "$PhysicalNames") {
            read_physical_names(in, m);  // text in both forms
        } else if (tok == "$Entities") {
            src.payload();
            long long counts[4];
            for (long long &c : counts) c = src.u64();
            for (int dim = 0; dim <= 3; dim++)
                for (long long i = 0; i < counts[dim]; i++) {
                    const int tag = (int)src.i32();
                    for (int k = 0; k < (dim == 0 ? 3 : 6); k++) src.f64();  // point, or bounding box
                    const long long nph = src.u64();
                    if (nph < 0 || nph > (1 << 20)) {
                        g_err = "malformed $Entities section";
                        return false;
                    }
                    std::vector<int> ph(nph);
                    for (auto &t : ph) t = (int)src.i32();
                    if (dim > 0) {
                        const long long nb = src.u64();
                        if (nb < 0 || nb > (1 << 20)) {
                            g_err = "malformed $Entities section";
                            return false;
                        }
                        for (long long k = 0; k < nb; k++) src.i32();
                    }
                    ent_phys[{dim, tag}] = ph;
                }
        } else if (tok == "$Nodes") {
            src.payload();
            const long long nblocks = src.u64(), nnodes = src.u64();
            src.u64();
            src.u64();  // min and max tag
            if (nnodes < 0 || nblocks < 0) {
                g_err = "malformed $Nodes section";
                return false;
            }
            m->pos.assign(3 * nnodes, 0.0);
            for (long long b = 0; b < nblocks; b++) {
                src.i32();  // entity dimension
                src.i32();  // entity tag
                const long long param = src.i32(), nb = src.u64();
                if (param) {
                    g_err = "parametric nodes are not supported";
                    return false;
                }
                if (nb < 0 || nb > nnodes) {
                    g_err = "malformed $Nodes section";
                    return false;
                }
                std::vector<long long> tags(nb);
                for (auto &t : tags) t = src.u64();
                for (long long k = 0; k < nb; k++) {
                    const double x = src.f64(), y = src.f64(), z = src.f64();
                    if (in.fail()) break;  // reported below
                    if (tags[k] < 1 || tags[k] > nnodes) {
                        g_err = "node tag out of range: " + std::to_string(tags[k]);
                        return false;
                    }
                    m->pos[3 * (tags[k] - 1)] = x;
                    m->pos[3 * (tags[k] - 1) + 1] = y;
                    m->pos[3 * (tags[k] - 1) + 2] = z;
                }
            }
        } else if (tok == "$Elements") {
            src.payload();
            const long long nblocks = src.u64();
            src.u64();  // number of elements
            src.u64();
            src.u64();  // min and max tag
            for (long long b = 0; b < nblocks; b++) {
                const int edim = (int)src.i32(), etag = (int)src.i32(), type = (int)src.i32();
                const long long nb = src.u64();
                const int nn = element_nodes(type);
                if (nn < 0) {
                    g_err = "unsupported element type " + std::to_string(type);
                    return false;
                }
                if (nb < 0) {
                    g_err = "malformed $Elements section";
                    return false;
                }
                const std::vector<int> &ph = ent_phys[{edim, etag}];
                for (long long k = 0; k < nb; k++) {
                    src.u64();  // element tag
                    int nodes[4];
                    for (int q = 0; q < nn; q++) nodes[q] = (int)src.u64();
                    for (int p : ph) {
                        if (type == 4) {
                            for (int q = 0; q < 4; q++) m->tets.push_back(nodes[q] - 1);
                            m->tet_phys.push_back(p);
                        } else if (type == 2) {
                            for (int q = 0; q < 3; q++) m->tris.push_back(nodes[q] - 1);
                            m->tri_phys.push_back(p);
                        }
                    }
                }
            }
        } else if (bin && tok.size() > 1 && tok[0] == '$' && tok.rfind("$End", 0) != 0) {
            skip_section(in, tok);
        }
        if (in.fail()) {
            g_err = "truncated or malformed gmsh file (" + tok + ")";
            return false;
        }
    }
    return true;
}

inline double mixed_product(const double *p0, const double *p1, const double *p2, const double *p3)
{
    double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double c[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
    double cx = b[1] * c[2] - b[2] * c[1], cy = b[2] * c[0] - b[0] * c[2], cz = b[0] * c[1] - b[1] * c[0];
    return a[0] * cx + a[1] * cy + a[2] * cz;
}

} // namespace

extern "C" {

const char *fgh_last_error(void) { return g_err.c_str(); }

int fgh_eval_expression(const char *expr, const char *vars, int64_t n, const double *vals, double *out)
{
    if (!expr || !vars || n < 0 || (n > 0 && !out)) {
        g_err = "fgh_eval_expression: bad arguments";
        return FG_ERR_ARG;
    }
    std::vector<std::string> names;
    for (const char *p = vars; *p;) {
        const char *e = std::strchr(p, ',');
        names.emplace_back(p, e ? (size_t)(e - p) : std::strlen(p));
        p = e ? e + 1 : p + std::strlen(p);
    }
    if (n > 0 && !names.empty() && !vals) {
        g_err = "fgh_eval_expression: bad arguments";
        return FG_ERR_ARG;
    }
    try {
        const fgh::Expr x = fgh::Expr::compile(expr, names);
        for (int64_t i = 0; i < n; i++) out[i] = x.eval(vals + (size_t)i * names.size());
    } catch (const fgh::ExprError &e) {
        g_err = e.what();
        return FG_ERR_ARG;
    }
    return FG_OK;
}

fgh_mesh *fgh_read_msh(const char *path, char *err, int errlen)
{
    std::ifstream in(path, std::ios::binary);
    auto fail = [&](const std::string &msg) -> fgh_mesh * {
        g_err = msg;
        if (err && errlen > 0) {
            std::strncpy(err, msg.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    };
    if (!in) return fail(std::string("cannot open ") + path);
    std::string tok;
    in >> tok;
    if (tok != "$MeshFormat") return fail("not a gmsh file: missing $MeshFormat");
    std::string version;
    int filetype, dsize;
    in >> version >> filetype >> dsize;
    if (filetype != 0 && filetype != 1) return fail("malformed $MeshFormat");
    const bool bin = filetype == 1;
    if (bin) {
        // a binary file follows the format line with the int 1 in the writer's byte order
        int32_t one = 0;
        in.get();
        in.read(reinterpret_cast<char *>(&one), sizeof one);
        if (one != 1) return fail("binary gmsh file of the other byte order");
        if (dsize != 8) return fail("binary gmsh file with data size " + std::to_string(dsize));
    }
    auto *m = new fgh_mesh();
    bool ok;
    if (version.rfind("2.", 0) == 0) ok = read_msh22(in, m, bin);
    else if (version.rfind("4.", 0) == 0) ok = read_msh41(in, m, bin);
    else {
        delete m;
        return fail("unsupported gmsh format " + version);
    }
    if (!ok) {
        delete m;
        return fail(g_err);
    }
    int64_t N = (int64_t)m->pos.size() / 3;
    for (int32_t v : m->tets)
        if (v < 0 || v >= N) {
            delete m;
            return fail("tetrahedron refers to a missing node");
        }
    for (int32_t v : m->tris)
        if (v < 0 || v >= N) {
            delete m;
            return fail("triangle refers to a missing node");
        }
    return m;
}

fgh_mesh *fgh_gen_cuboid(double x0, double y0, double z0, double lx, double ly, double lz, int nx,
                         int ny, int nz, double jitter, uint64_t seed)
{
    if (nx < 1 || ny < 1 || nz < 1) {
        g_err = "cuboid needs at least one cell per axis";
        return nullptr;
    }
    auto *m = new fgh_mesh();
    const double dx = lx / nx, dy = ly / ny, dz = lz / nz;
    const long NN = (long)(nx + 1) * (ny + 1) * (nz + 1);
    m->pos.resize(3 * NN);
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    auto idx = [&](int i, int j, int k) { return (int32_t)((long)(nz + 1) * (ny + 1) * i + (long)(nz + 1) * j + k); };
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++)
            for (int k = 0; k <= nz; k++) {
                double jx = U(gen), jy = U(gen), jz = U(gen);
                double *p = &m->pos[3 * (long)idx(i, j, k)];
                p[0] = x0 + i * dx + ((i > 0 && i < nx) ? jitter * dx * jx : 0.0);
                p[1] = y0 + j * dy + ((j > 0 && j < ny) ? jitter * dy * jy : 0.0);
                p[2] = z0 + k * dz + ((k > 0 && k < nz) ? jitter * dz * jz : 0.0);
            }
    m->tets.reserve(24L * nx * ny * nz);
    auto tri = [&](int a, int b, int c) {
        m->tris.push_back(a);
        m->tris.push_back(b);
        m->tris.push_back(c);
        m->tri_phys.push_back(200);
    };
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                int A = idx(i, j, k), B = idx(i + 1, j, k), C = idx(i + 1, j + 1, k), D = idx(i, j + 1, k);
                int E = idx(i, j, k + 1), F = idx(i + 1, j, k + 1), G = idx(i + 1, j + 1, k + 1), H = idx(i, j + 1, k + 1);
                const int T[6][4] = {{A, B, D, H}, {A, B, E, H}, {E, F, H, B}, {B, C, D, H}, {B, C, G, H}, {F, B, G, H}};
                for (auto &t : T) {
                    for (int q = 0; q < 4; q++) m->tets.push_back(t[q]);
                    m->tet_phys.push_back(300);
                }
                if (k == 0) { tri(A, B, D); tri(B, C, D); }
                if (k == nz - 1) { tri(E, F, H); tri(F, G, H); }
                if (i == 0) { tri(A, D, H); tri(A, E, H); }
                if (i == nx - 1) { tri(B, C, G); tri(B, F, G); }
                if (j == 0) { tri(A, B, E); tri(B, E, F); }
                if (j == ny - 1) { tri(C, G, H); tri(H, D, C); }
            }
    m->phys_names[{2, 200}] = "surface";
    m->phys_names[{3, 300}] = "volume";
    return m;
}

fgh_mesh *fgh_gen_ellipsoid(double cx, double cy, double cz, double a, double b, double c, double h)
{
    if (!(a > 0 && b > 0 && c > 0 && h > 0)) {
        g_err = "ellipsoid needs positive semi-axes and spacing";
        return nullptr;
    }
    /* grid covering the box, cells kept when their centre is inside */
    const int nx = std::max(2, (int)std::ceil(2 * a / h)), ny = std::max(2, (int)std::ceil(2 * b / h)),
              nz = std::max(2, (int)std::ceil(2 * c / h));
    const double dx = 2 * a / nx, dy = 2 * b / ny, dz = 2 * c / nz;
    auto inside = [&](double x, double y, double z) {
        double X = (x - cx) / a, Y = (y - cy) / b, Z = (z - cz) / c;
        return X * X + Y * Y + Z * Z <= 1.0;
    };
    auto gidx = [&](int i, int j, int k) { return (long)(nz + 1) * (ny + 1) * i + (long)(nz + 1) * j + k; };
    std::vector<int32_t> remap((size_t)(nx + 1) * (ny + 1) * (nz + 1), -1);
    auto *m = new fgh_mesh();
    std::vector<char> boundary;
    auto node = [&](int i, int j, int k) {
        long g = gidx(i, j, k);
        if (remap[g] < 0) {
            remap[g] = (int32_t)(m->pos.size() / 3);
            m->pos.push_back(cx - a + i * dx);
            m->pos.push_back(cy - b + j * dy);
            m->pos.push_back(cz - c + k * dz);
        }
        return remap[g];
    };
    std::vector<char> cell((size_t)nx * ny * nz, 0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                cell[((size_t)i * ny + j) * nz + k] =
                    inside(cx - a + (i + 0.5) * dx, cy - b + (j + 0.5) * dy, cz - c + (k + 0.5) * dz);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                if (!cell[((size_t)i * ny + j) * nz + k]) continue;
                int A = node(i, j, k), B = node(i + 1, j, k), C = node(i + 1, j + 1, k), D = node(i, j + 1, k);
                int E = node(i, j, k + 1), F = node(i + 1, j, k + 1), G = node(i + 1, j + 1, k + 1), H = node(i, j + 1, k + 1);
                const int T[6][4] = {{A, B, D, H}, {A, B, E, H}, {E, F, H, B}, {B, C, D, H}, {B, C, G, H}, {F, B, G, H}};
                for (auto &t : T) {
                    for (int q = 0; q < 4; q++) m->tets.push_back(t[q]);
                    m->tet_phys.push_back(300);
                }
            }
    /* nodes on the staircase boundary are moved half-way towards the ellipsoid surface along the
     * ray from the centre, which smooths the surface while keeping every tetrahedron valid */
    std::map<std::array<int, 3>, int> facecount;
    for (size_t t = 0; t < m->tets.size() / 4; t++) {
        const int32_t *q = &m->tets[4 * t];
        const int F[4][3] = {{q[0], q[2], q[1]}, {q[1], q[2], q[3]}, {q[2], q[0], q[3]}, {q[3], q[0], q[1]}};
        for (auto &f : F) {
            std::array<int, 3> s = {f[0], f[1], f[2]};
            std::sort(s.begin(), s.end());
            facecount[s]++;
        }
    }
    boundary.assign(m->pos.size() / 3, 0);
    for (auto &kv : facecount)
        if (kv.second == 1)
            for (int v : kv.first) boundary[v] = 1;
    for (size_t v = 0; v < boundary.size(); v++) {
        if (!boundary[v]) continue;
        double *p = &m->pos[3 * v];
        double X = (p[0] - cx) / a, Y = (p[1] - cy) / b, Z = (p[2] - cz) / c;
        double r = std::sqrt(X * X + Y * Y + Z * Z);
        if (r <= 0) continue;
        double s = 0.5 * (1.0 + 1.0 / r);
        p[0] = cx + (p[0] - cx) * s;
        p[1] = cy + (p[1] - cy) * s;
        p[2] = cz + (p[2] - cz) * s;
    }
    m->phys_names[{3, 300}] = "volume";
    return m;
}

fgh_mesh *fgh_gen_cylinder(double cx, double cy, double z0, double r, double length, double h)
{
    if (!(r > 0 && length > 0 && h > 0)) {
        g_err = "cylinder needs a positive radius, length and spacing";
        return nullptr;
    }
    /* grid over the bounding box, cells kept when their centre is inside the disc */
    const int nx = std::max(2, (int)std::ceil(2 * r / h)), nz = std::max(1, (int)std::ceil(length / h));
    const int ny = nx;
    const double dx = 2 * r / nx, dy = dx, dz = length / nz;
    const double x0 = cx - r, y0 = cy - r;
    std::vector<char> col((size_t)nx * ny, 0);  // the kept columns of cells (the same at every z)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const double X = x0 + (i + 0.5) * dx - cx, Y = y0 + (j + 0.5) * dy - cy;
            col[(size_t)i * ny + j] = X * X + Y * Y <= r * r;
        }
    auto kept = [&](int i, int j) { return i >= 0 && j >= 0 && i < nx && j < ny && col[(size_t)i * ny + j]; };
    /* nodes of kept columns: the (i, j) grid points with a kept cell around them, all z levels */
    std::vector<int32_t> pid((size_t)(nx + 1) * (ny + 1), -1);
    int np = 0;
    std::vector<char> lateral;  // grid point on the lateral boundary (a missing column around it)
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++) {
            int in = 0, tot = 0;
            for (int a = -1; a <= 0; a++)
                for (int b = -1; b <= 0; b++) {
                    tot++;
                    in += kept(i + a, j + b);
                }
            if (in == 0) continue;
            pid[(size_t)i * (ny + 1) + j] = np++;
            lateral.push_back(in < tot);
        }
    auto *m = new fgh_mesh();
    m->pos.resize((size_t)np * (nz + 1) * 3);
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++) {
            const int p = pid[(size_t)i * (ny + 1) + j];
            if (p < 0) continue;
            double X = x0 + i * dx - cx, Y = y0 + j * dy - cy;
            if (lateral[p]) {
                /* staircase rim: half-way towards the cylinder surface along the radius */
                const double rho = std::sqrt(X * X + Y * Y);
                if (rho > 0) {
                    const double s = 0.5 * (1.0 + r / rho);
                    X *= s;
                    Y *= s;
                }
            }
            for (int k = 0; k <= nz; k++) {
                double *q = &m->pos[3 * ((size_t)k * np + p)];
                q[0] = cx + X;
                q[1] = cy + Y;
                q[2] = z0 + k * dz;
            }
        }
    auto node = [&](int i, int j, int k) { return (int32_t)((size_t)k * np + pid[(size_t)i * (ny + 1) + j]); };
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                if (!kept(i, j)) continue;
                int A = node(i, j, k), B = node(i + 1, j, k), C = node(i + 1, j + 1, k), D = node(i, j + 1, k);
                int E = node(i, j, k + 1), F = node(i + 1, j, k + 1), G = node(i + 1, j + 1, k + 1),
                    H = node(i, j + 1, k + 1);
                const int T[6][4] = {{A, B, D, H}, {A, B, E, H}, {E, F, H, B}, {B, C, D, H}, {B, C, G, H}, {F, B, G, H}};
                for (auto &t : T) {
                    for (int q = 0; q < 4; q++) m->tets.push_back(t[q]);
                    m->tet_phys.push_back(300);
                }
            }
    m->phys_names[{3, 300}] = "volume";
    return m;
}

int fgh_append(fgh_mesh *dst, const fgh_mesh *src)
{
    if (!dst || !src || dst->prepared) return FG_ERR_ARG;
    int32_t off = (int32_t)(dst->pos.size() / 3);
    dst->pos.insert(dst->pos.end(), src->pos.begin(), src->pos.end());
    for (int32_t v : src->tets) dst->tets.push_back(v + off);
    dst->tet_phys.insert(dst->tet_phys.end(), src->tet_phys.begin(), src->tet_phys.end());
    for (int32_t v : src->tris) dst->tris.push_back(v + off);
    dst->tri_phys.insert(dst->tri_phys.end(), src->tri_phys.begin(), src->tri_phys.end());
    for (auto &kv : src->phys_names) dst->phys_names.insert(kv);
    return FG_OK;
}

void fgh_free(fgh_mesh *m) { delete m; }

int fgh_raw_counts(const fgh_mesh *m, int64_t *n_nodes, int64_t *n_tets, int64_t *n_tris)
{
    if (!m) return FG_ERR_ARG;
    if (n_nodes) *n_nodes = (int64_t)m->pos.size() / 3;
    if (n_tets) *n_tets = (int64_t)m->tet_phys.size();
    if (n_tris) *n_tris = (int64_t)m->tri_phys.size();
    return FG_OK;
}

int fgh_write_msh22(const fgh_mesh *m, const char *path)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return FG_ERR_ARG;
    std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n%zu\n", m->phys_names.size());
    for (auto &kv : m->phys_names) std::fprintf(f, "%d %d \"%s\"\n", kv.first.first, kv.first.second, kv.second.c_str());
    size_t N = m->pos.size() / 3;
    std::fprintf(f, "$EndPhysicalNames\n$Nodes\n%zu\n", N);
    for (size_t i = 0; i < N; i++)
        std::fprintf(f, "%zu %.17g %.17g %.17g\n", i + 1, m->pos[3 * i], m->pos[3 * i + 1], m->pos[3 * i + 2]);
    size_t T = m->tet_phys.size(), F = m->tri_phys.size();
    std::fprintf(f, "$EndNodes\n$Elements\n%zu\n", T + F);
    size_t id = 1;
    for (size_t t = 0; t < T; t++, id++)
        std::fprintf(f, "%zu 4 2 %d 1 %d %d %d %d\n", id, m->tet_phys[t], m->tets[4 * t] + 1, m->tets[4 * t + 1] + 1,
                     m->tets[4 * t + 2] + 1, m->tets[4 * t + 3] + 1);
    for (size_t t = 0; t < F; t++, id++)
        std::fprintf(f, "%zu 2 2 %d 1 %d %d %d\n", id, m->tri_phys[t], m->tris[3 * t] + 1, m->tris[3 * t + 1] + 1,
                     m->tris[3 * t + 2] + 1);
    std::fprintf(f, "$EndElements\n");
    std::fclose(f);
    return FG_OK;
}

int fgh_prepare(fgh_mesh *m, double scale, int n_vol, const char *const *vol_names,
                const fg_volume_region *vol, int n_surf, const char *const *surf_names,
                const fg_surface_region *surf)
{
    if (!m || n_vol < 1 || !vol_names || !vol || scale <= 0) {
        g_err = "fgh_prepare: invalid arguments";
        return FG_ERR_ARG;
    }
    /* named regions must exist in the file (checkMeshFile, src/read.cpp:35-71) */
    auto find_tags = [&](int dim, const char *name) {
        std::vector<int> tags;
        for (auto &kv : m->phys_names)
            if (kv.first.first == dim && kv.second == name) tags.push_back(kv.first.second);
        return tags;
    };
    std::map<int, int> vol_of_tag, surf_of_tag;
    for (int r = 0; r < n_vol; r++) {
        auto tags = find_tags(3, vol_names[r]);
        if (tags.empty()) {
            g_err = std::string("volume region ") + vol_names[r] + " not found in the mesh";
            return FG_ERR_MESH;
        }
        for (int t : tags) vol_of_tag[t] = r;
    }
    for (int r = 0; r < n_surf; r++) {
        auto tags = find_tags(2, surf_names[r]);
        if (tags.empty()) {
            g_err = std::string("surface region ") + surf_names[r] + " not found in the mesh";
            return FG_ERR_MESH;
        }
        for (int t : tags) surf_of_tag[t] = r;
    }
    const size_t N = m->pos.size() / 3;
    std::vector<double> P(m->pos.size());
    for (size_t i = 0; i < P.size(); i++) P[i] = m->pos[i] * scale;
    /* tetrahedra of named regions, positively oriented */
    std::vector<int32_t> tets, treg;
    for (size_t t = 0; t < m->tet_phys.size(); t++) {
        auto it = vol_of_tag.find(m->tet_phys[t]);
        if (it == vol_of_tag.end()) continue;
        int32_t q[4] = {m->tets[4 * t], m->tets[4 * t + 1], m->tets[4 * t + 2], m->tets[4 * t + 3]};
        double mp = mixed_product(&P[3 * q[0]], &P[3 * q[1]], &P[3 * q[2]], &P[3 * q[3]]);
        if (std::fabs(mp) < 1e-40) {
            g_err = "singular tetrahedron, mixed product " + std::to_string(mp);
            return FG_ERR_MESH;
        }
        if (mp < 0) std::swap(q[2], q[3]);
        tets.insert(tets.end(), q, q + 4);
        treg.push_back(it->second);
    }
    if (tets.empty()) {
        g_err = "not a single tetrahedron is present in the mesh";
        return FG_ERR_MESH;
    }
    std::vector<int32_t> tris, freg;
    for (size_t t = 0; t < m->tri_phys.size(); t++) {
        auto it = surf_of_tag.find(m->tri_phys[t]);
        if (it == surf_of_tag.end()) continue;
        tris.insert(tris.end(), &m->tris[3 * t], &m->tris[3 * t] + 3);
        freg.push_back(it->second);
    }
    /* boundary and interface triangles */
    std::vector<Face> faces;
    faces.reserve(4 * treg.size() + freg.size());
    for (size_t t = 0; t < treg.size(); t++) {
        const int32_t *q = &tets[4 * t];
        faces.push_back(make_face(q[0], q[2], q[1], treg[t], false));
        faces.push_back(make_face(q[1], q[2], q[3], treg[t], false));
        faces.push_back(make_face(q[2], q[0], q[3], treg[t], false));
        faces.push_back(make_face(q[3], q[0], q[1], treg[t], false));
    }
    for (size_t t = 0; t < freg.size(); t++)
        faces.push_back(make_face(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], freg[t], true));
    std::sort(faces.begin(), faces.end());
    const int added_region = n_surf;
    for (size_t lo = 0; lo < faces.size();) {
        size_t hi = lo + 1;
        while (hi < faces.size() && faces[hi].n == faces[lo].n) hi++;
        int nsurf = 0, ntet = 0, r0 = -1, r1 = -1;
        for (size_t k = lo; k < hi; k++) {
            if (faces[k].surface) nsurf++;
            else {
                if (ntet == 0) r0 = faces[k].region;
                else r1 = faces[k].region;
                ntet++;
            }
        }
        if (nsurf > 1) {
            g_err = "bad mesh: " + std::to_string(nsurf) + " instances of the same surface triangle";
            return FG_ERR_MESH;
        }
        if (ntet == 0) {
            g_err = "bad mesh: a surface triangle belongs to no tetrahedron";
            return FG_ERR_MESH;
        }
        if (ntet > 2) {
            g_err = "bad mesh: a triangular face shared by " + std::to_string(ntet) + " tetrahedra";
            return FG_ERR_MESH;
        }
        if (ntet == 2 && nsurf == 1 && r0 == r1) {
            g_err = "bad mesh: an internal triangle is in a surface region";
            return FG_ERR_MESH;
        }
        if (nsurf == 0 && (ntet == 1 || r0 != r1)) {
            tris.insert(tris.end(), faces[lo].n.begin(), faces[lo].n.end());
            freg.push_back(added_region);
        }
        lo = hi;
    }
    /* signed Ms jump of every triangle: + Ms of each tet whose outward face has the triangle's
     * orientation, - otherwise */
    std::vector<double> dMs(freg.size(), 0.0);
    for (size_t t = 0; t < freg.size(); t++) {
        Face f = make_face(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], 0, true);
        auto it = std::lower_bound(faces.begin(), faces.end(), f);
        for (; it != faces.end() && it->n == f.n; ++it)
            if (!it->surface) {
                double Ms = vol[it->region].Ms;
                dMs[t] += (f.flipped ^ it->flipped) ? -Ms : Ms;
            }
    }
    /* bounding box (doOnNodes, src/mesh.cpp:108-119): max starts from DBL_MIN as the reference */
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {DBL_MIN, DBL_MIN, DBL_MIN};
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            mn[k] = std::min(mn[k], P[3 * i + k]);
            mx[k] = std::max(mx[k], P[3 * i + k]);
        }
    for (int k = 0; k < 3; k++) {
        m->l[k] = mx[k] - mn[k];
        m->c[k] = 0.5 * (mx[k] + mn[k]);
    }
    int axis;
    if (m->l[0] > m->l[1]) axis = m->l[0] > m->l[2] ? 0 : 2;
    else axis = m->l[1] > m->l[2] ? 1 : 2;
    std::vector<int32_t> perm(N);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](int a, int b) { return P[3 * a + axis] < P[3 * b + axis]; });
    m->node_index.assign(N, 0);
    for (size_t i = 0; i < N; i++) m->node_index[perm[i]] = (int32_t)i;
    m->ppos.resize(3 * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) m->ppos[3 * i + k] = P[3 * perm[i] + k];
    for (auto &v : tets) v = m->node_index[v];
    for (auto &v : tris) v = m->node_index[v];
    m->ptets = std::move(tets);
    m->ptet_reg = std::move(treg);
    m->ptris = std::move(tris);
    m->ptri_reg = std::move(freg);
    m->pdMs = std::move(dMs);
    m->vol.assign(vol, vol + n_vol);
    m->surf.assign(n_surf + 1, fg_surface_region{});
    for (int r = 0; r < n_surf; r++) m->surf[r] = surf[r];
    m->surf[n_surf].Ks = 0;
    m->surf[n_surf].uk[2] = 1;
    m->surf[n_surf].suppress_charges = 0;
    m->prepared = true;
    return FG_OK;
}

int fgh_view(const fgh_mesh *m, fg_mesh *out)
{
    if (!m || !out || !m->prepared) return FG_ERR_STATE;
    std::memset(out, 0, sizeof(*out));
    out->n_nodes = (int32_t)(m->ppos.size() / 3);
    out->n_tets = (int32_t)m->ptet_reg.size();
    out->n_tris = (int32_t)m->ptri_reg.size();
    out->n_vol_regions = (int32_t)m->vol.size();
    out->n_surf_regions = (int32_t)m->surf.size();
    out->pos = m->ppos.data();
    out->tet_nodes = m->ptets.data();
    out->tet_region = m->ptet_reg.data();
    out->tri_nodes = m->ptris.data();
    out->tri_region = m->ptri_reg.data();
    out->tri_dMs = m->pdMs.data();
    out->vol = m->vol.data();
    out->surf = m->surf.data();
    return FG_OK;
}

const int32_t *fgh_node_index(const fgh_mesh *m) { return (m && m->prepared) ? m->node_index.data() : nullptr; }

int fgh_bbox(const fgh_mesh *m, double l[3], double c[3])
{
    if (!m || !m->prepared) return FG_ERR_STATE;
    for (int k = 0; k < 3; k++) {
        l[k] = m->l[k];
        c[k] = m->c[k];
    }
    return FG_OK;
}

/* Mesh::mesh::readSol (src/read.cpp:199-249): u and phi in file order, mapped through
 * node_index; outputs are in sorted order. */
int fgh_read_sol(const fgh_mesh *m, const char *path, double *u, double *phi, double *t)
{
    if (!m || !m->prepared) return FG_ERR_STATE;
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        g_err = std::string("cannot open file ") + path;
        return FG_ERR_ARG;
    }
    bool flag_t = false;
    std::string line;
    while (in.peek() == '#' || in.peek() == '\n') {
        std::getline(in, line);
        auto pos = line.find("## time:");
        if (pos != std::string::npos) {
            *t = std::stod(line.substr(pos + 8));
            flag_t = true;
        }
    }
    if (!flag_t) {
        g_err = "no ## time: tag in input .sol file";
        return FG_ERR_ARG;
    }
    const size_t N = m->node_index.size();
    for (size_t i = 0; i < N; i++) {
        unsigned long idx;
        double mx, my, mz, ph;
        if (!(in >> idx >> mx >> my >> mz >> ph)) {
            g_err = "truncated .sol file";
            return FG_ERR_ARG;
        }
        if (idx != i) {
            g_err = "mesh node index mismatch between mesh and input .sol file";
            return FG_ERR_ARG;
        }
        int32_t j = m->node_index[i];
        u[3 * j] = mx;
        u[3 * j + 1] = my;
        u[3 * j + 2] = mz;
        if (phi) phi[j] = ph;
    }
    return FG_OK;
}

/* Mesh::mesh::savesol (src/save.cpp:164-186) */
int fgh_write_sol(const fgh_mesh *m, const char *path, const char *metadata, int precision,
                  const double *u, const double *phi)
{
    return fgh_write_sol_s(m, path, metadata, precision, u, phi, nullptr);
}

int fgh_write_sol_s(const fgh_mesh *m, const char *path, const char *metadata, int precision,
                    const double *u, const double *phi, const double *s)
{
    if (!m || !m->prepared) return FG_ERR_STATE;
    FILE *f = std::fopen(path, "w");
    if (!f) {
        g_err = std::string("cannot open file ") + path;
        return FG_ERR_ARG;
    }
    if (metadata) std::fputs(metadata, f);
    const size_t N = m->node_index.size();
    for (size_t i = 0; i < N; i++) {
        int32_t j = m->node_index[i];
        std::fprintf(f, "%zu\t%.*e\t%.*e\t%.*e\t%.*e", i, precision, u[3 * j], precision, u[3 * j + 1], precision,
                     u[3 * j + 2], precision, phi ? phi[j] : 0.0);
        if (s)
            std::fprintf(f, "\t%.*e\t%.*e\t%.*e", precision, s[3 * j], precision, s[3 * j + 1], precision,
                         s[3 * j + 2]);
        std::fputc('\n', f);
    }
    std::fclose(f);
    return FG_OK;
}

} // extern "C"
