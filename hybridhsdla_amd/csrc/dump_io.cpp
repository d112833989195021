// Fleur dump reader and H/S writer (HDF5 C API), see include/hsgen_io.h.
#include "hsgen_io.h"

#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

hsgen_status fail(hsgen_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

// owns one HDF5 identifier
struct H {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H() = default;
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    H(const H&) = delete;
    H& operator=(const H&) = delete;
    H(H&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    ~H()
    {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
    bool ok() const { return id >= 0; }
};

void quiet()
{
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

hid_t cx_type()   // complex128 as array[2] of double, as the files store it
{
    const hsize_t d2[1] = {2};
    return H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, d2);
}

bool read_attr_int(hid_t obj, const char* name, int32_t* v)
{
    if (H5Aexists(obj, name) <= 0) return false;
    H a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Aread(a, H5T_NATIVE_INT32, v) >= 0;
}

bool write_attr_int(hid_t obj, const char* name, int32_t v)
{
    const hsize_t one[1] = {1};
    H sp(H5Screate_simple(1, one, nullptr), H5Sclose);
    H a(H5Acreate2(obj, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a, H5T_NATIVE_INT32, &v) >= 0;
}

// element count of a dataset's dataspace
int64_t extent(hid_t ds, std::vector<hsize_t>* dims = nullptr)
{
    H sp(H5Dget_space(ds), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(sp);
    if (nd < 0) return -1;
    std::vector<hsize_t> d(nd > 0 ? nd : 1, 1);
    if (nd > 0) H5Sget_simple_extent_dims(sp, d.data(), nullptr);
    int64_t n = 1;
    for (int i = 0; i < nd; ++i) n *= (int64_t)d[i];
    if (dims) *dims = d;
    return n;
}

bool read_doubles(hid_t file, const char* name, std::vector<double>& out, std::vector<hsize_t>* dims = nullptr)
{
    H ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return false;
    const int64_t n = extent(ds, dims);
    if (n < 0) return false;
    out.assign((size_t)n, 0.0);
    return H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

bool read_ints(hid_t file, const char* name, std::vector<int32_t>& out)
{
    H ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return false;
    const int64_t n = extent(ds);
    if (n < 0) return false;
    out.assign((size_t)n, 0);
    return H5Dread(ds, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

// complex dataset whose element type is array[2] of double (possibly inside
// further array dimensions): read everything as a flat double array
bool read_complex(hid_t loc, const char* name, std::vector<double>& out, int64_t* nelem = nullptr)
{
    H ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return false;
    H ft(H5Dget_type(ds), H5Tclose);
    const size_t esize = H5Tget_size(ft);
    const int64_t n = extent(ds);
    if (n < 0 || esize % 16 != 0) return false;
    const int64_t per = (int64_t)(esize / 16);            // complex values per dataset element
    // memory type with the file type's shape, native doubles
    H mt(H5Tget_native_type(ft, H5T_DIR_ASCEND), H5Tclose);
    out.assign((size_t)(2 * n * per), 0.0);
    if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0) return false;
    if (nelem) *nelem = n * per;
    return true;
}

double det3(const double* m)   // column-major 3x3
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}

void matmul3(const double* a, const double* b, double* c)   // column-major c = a b
{
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i + 3 * k] * b[k + 3 * j];
            c[i + 3 * j] = s;
        }
}

struct AtomRec {
    int32_t type;
    double pos[3];
    int32_t rot[9];
};
struct MeshRec {
    int32_t num;
    double dx;
    double rmt;
};

}  // namespace

struct hsgen_dump {
    hsgen_dump_info info{};
    std::vector<int32_t> lmax_t, lnonsph_t;       // per type
    std::vector<MeshRec> mesh;                    // per type
    std::vector<double> energies;                 // types x (max_lmax+1)
    std::vector<double> potential;                // types x max_mesh_pts
    std::vector<std::vector<double>> tmat;        // per type: 4 n n complex (interleaved)
    std::vector<AtomRec> atoms;
    std::vector<int32_t> num_gpts;
    std::vector<double> kpts;                     // 3 per k
    std::vector<int32_t> gmesh;                   // max_gpts x 3 per k
    std::vector<std::vector<double>> hs;          // per k: H packed, S packed (2 entries)
    // final data from /hsgen, when present
    std::vector<double> f_match, f_rot;
    std::vector<std::vector<double>> f_t;         // per atom: aa, ab, bb (3 n n complex)
    // system description built on demand
    bool built = false;
    std::vector<int32_t> s_lmax;
    std::vector<double> s_rmt, s_pos, s_rot, s_match;
    std::vector<std::vector<double>> s_t;         // per type (or atom): aa, ab, bb blocks
    std::vector<const double*> s_taa, s_tab, s_tbb;
};

struct hsgen_store {
    hid_t file = -1, gh = -1, gs = -1;
};

extern "C" {

const char* hsgen_io_last_error(void)
{
    return g_err.c_str();
}

hsgen_status hsgen_dump_open(const char* path, hsgen_dump** out)
{
    if (!path || !out) return fail(HSGEN_ERR_ARG, "null argument");
    *out = nullptr;
    quiet();
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(HSGEN_ERR_ARG, "cannot open HDF5 file %s", path);
    hsgen_dump* d = new hsgen_dump();
    auto bad = [&](const char* what) {
        delete d;
        return fail(HSGEN_ERR_ARG, "%s: cannot read %s", path, what);
    };
    // basis (FleurData.cpp:580-595): 9 doubles, taken as column-major
    std::vector<double> v;
    if (!read_doubles(f, "/real_basis", v) || v.size() != 9) return bad("/real_basis");
    std::memcpy(d->info.real_basis, v.data(), sizeof d->info.real_basis);
    if (!read_doubles(f, "/rec_basis", v) || v.size() != 9) return bad("/rec_basis");
    std::memcpy(d->info.rec_basis, v.data(), sizeof d->info.rec_basis);
    d->info.omega = det3(d->info.real_basis);
    // atoms (FleurData.cpp:375-432)
    {
        H ds(H5Dopen2(f, "/atoms", H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return bad("/atoms");
        if (!read_attr_int(ds, "num_atoms", &d->info.num_atoms) || !read_attr_int(ds, "num_types", &d->info.num_types))
            return bad("/atoms attributes");
        if (extent(ds) != d->info.num_atoms || d->info.num_atoms <= 0 || d->info.num_types <= 0)
            return bad("/atoms extent");
        H mt(H5Tcreate(H5T_COMPOUND, sizeof(AtomRec)), H5Tclose);
        const hsize_t d3[1] = {3}, d33[2] = {3, 3};
        H a3(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, d3), H5Tclose);
        H a33(H5Tarray_create2(H5T_NATIVE_INT32, 2, d33), H5Tclose);
        H5Tinsert(mt, "type", HOFFSET(AtomRec, type), H5T_NATIVE_INT32);
        H5Tinsert(mt, "position", HOFFSET(AtomRec, pos), a3);
        H5Tinsert(mt, "rotation", HOFFSET(AtomRec, rot), a33);
        d->atoms.resize(d->info.num_atoms);
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, d->atoms.data()) < 0) return bad("/atoms records");
        for (const AtomRec& a : d->atoms)
            if (a.type < 1 || a.type > d->info.num_types) return bad("/atoms type index");
    }
    const int nt = d->info.num_types;
    // radial meshes (FleurData.cpp:355-369)
    {
        H ds(H5Dopen2(f, "/mesh", H5P_DEFAULT), H5Dclose);
        if (!ds.ok() || extent(ds) != nt) return bad("/mesh");
        H mt(H5Tcreate(H5T_COMPOUND, sizeof(MeshRec)), H5Tclose);
        H5Tinsert(mt, "num_points", HOFFSET(MeshRec, num), H5T_NATIVE_INT32);
        H5Tinsert(mt, "dx", HOFFSET(MeshRec, dx), H5T_NATIVE_DOUBLE);
        H5Tinsert(mt, "rmt", HOFFSET(MeshRec, rmt), H5T_NATIVE_DOUBLE);
        d->mesh.resize(nt);
        if (H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, d->mesh.data()) < 0) return bad("/mesh records");
        if (!read_attr_int(ds, "max_mesh_pts", &d->info.max_mesh_pts)) return bad("/mesh max_mesh_pts");
    }
    // cutoffs (FleurData.cpp:326-348)
    if (!read_ints(f, "/lmax", d->lmax_t) || (int)d->lmax_t.size() != nt) return bad("/lmax");
    if (!read_ints(f, "/lmax_nonspherical", d->lnonsph_t) || (int)d->lnonsph_t.size() != nt)
        return bad("/lmax_nonspherical");
    d->info.max_lmax = 0;
    for (int t = 0; t < nt; ++t) {
        if (d->lmax_t[t] < 0 || d->lnonsph_t[t] < 0 || d->lnonsph_t[t] > d->lmax_t[t]) return bad("cutoffs");
        d->info.max_lmax = std::max(d->info.max_lmax, d->lmax_t[t]);
    }
    // energy parameters: types x (max_lmax+1) (FleurData.cpp:309-318)
    std::vector<hsize_t> dims;
    if (!read_doubles(f, "/energy_params", d->energies, &dims) || dims.size() != 2 || (int)dims[0] != nt ||
        (int)dims[1] < d->info.max_lmax + 1)
        return bad("/energy_params");
    const int ne = (int)dims[1];
    if (ne != d->info.max_lmax + 1) {   // keep max_lmax+1 per type
        std::vector<double> e2((size_t)nt * (d->info.max_lmax + 1));
        for (int t = 0; t < nt; ++t)
            for (int l = 0; l <= d->info.max_lmax; ++l) e2[(size_t)t * (d->info.max_lmax + 1) + l] = d->energies[(size_t)t * ne + l];
        d->energies.swap(e2);
    }
    // potential: types x max_mesh_pts (FleurData.cpp:293-305)
    if (!read_doubles(f, "/potential", d->potential, &dims) || dims.size() != 2 || (int)dims[0] != nt ||
        (int)dims[1] != d->info.max_mesh_pts)
        return bad("/potential");
    for (int t = 0; t < nt; ++t)
        if (d->mesh[t].num > d->info.max_mesh_pts) return bad("/mesh num_points");
    // k-points (FleurData.cpp:239-290)
    {
        H ds(H5Dopen2(f, "/k_meshes", H5P_DEFAULT), H5Dclose);
        if (!ds.ok()) return bad("/k_meshes");
        if (!read_attr_int(ds, "num_kpts", &d->info.num_kpts) || !read_attr_int(ds, "max_gpts", &d->info.max_gpts))
            return bad("/k_meshes attributes");
        const int nk = d->info.num_kpts, mg = d->info.max_gpts;
        if (nk < 0 || mg < 0 || extent(ds) != nk) return bad("/k_meshes extent");
        const size_t rec = sizeof(int32_t) * 2 + sizeof(double) * 3 + sizeof(int32_t) * 3 * (size_t)mg;
        std::vector<char> buf(rec * (size_t)std::max(nk, 1));
        H mt(H5Tcreate(H5T_COMPOUND, rec), H5Tclose);
        const hsize_t d3[1] = {3}, dg[2] = {(hsize_t)mg, 3};
        H a3(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, d3), H5Tclose);
        H ag(H5Tarray_create2(H5T_NATIVE_INT32, 2, dg), H5Tclose);
        const size_t o_k = 8, o_g = 8 + 24;
        H5Tinsert(mt, "num_gpts", 0, H5T_NATIVE_INT32);
        H5Tinsert(mt, "k_point", o_k, a3);
        if (mg > 0) H5Tinsert(mt, "g_mesh", o_g, ag);
        if (nk > 0 && H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) return bad("/k_meshes records");
        d->num_gpts.resize(nk);
        d->kpts.resize(3 * (size_t)nk);
        d->gmesh.resize(3 * (size_t)mg * nk);
        for (int k = 0; k < nk; ++k) {
            const char* r = buf.data() + rec * k;
            std::memcpy(&d->num_gpts[k], r, sizeof(int32_t));
            std::memcpy(&d->kpts[3 * (size_t)k], r + o_k, 3 * sizeof(double));
            std::memcpy(&d->gmesh[3 * (size_t)mg * k], r + o_g, 3 * sizeof(int32_t) * (size_t)mg);
            if (d->num_gpts[k] < 0 || d->num_gpts[k] > mg) return bad("/k_meshes num_gpts");
        }
    }
    // T blocks (FleurData.cpp:216-236): /tmat/<type> holds 4 n x n complex
    {
        H g(H5Gopen2(f, "/tmat", H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return bad("/tmat");
        d->tmat.resize(nt);
        for (int t = 0; t < nt; ++t) {
            const std::string nm = std::to_string(t + 1);
            int64_t nel = 0;
            if (!read_complex(g, nm.c_str(), d->tmat[t], &nel)) return bad("/tmat/<type>");
            const int64_t n = (int64_t)(d->lmax_t[t] + 1) * (d->lmax_t[t] + 1);
            if (nel != 4 * n * n) return bad("/tmat/<type> size");
        }
    }
    // Fleur's H and S, if present (binmat.f90:194-219)
    if (H5Lexists(f, "/matrix_hamilton", H5P_DEFAULT) > 0 && H5Lexists(f, "/matrix_overlap", H5P_DEFAULT) > 0) {
        H gh(H5Gopen2(f, "/matrix_hamilton", H5P_DEFAULT), H5Gclose);
        H gs(H5Gopen2(f, "/matrix_overlap", H5P_DEFAULT), H5Gclose);
        for (int k = 0;; ++k) {
            const std::string nm = std::to_string(k + 1);
            if (H5Lexists(gh, nm.c_str(), H5P_DEFAULT) <= 0 || H5Lexists(gs, nm.c_str(), H5P_DEFAULT) <= 0) break;
            std::vector<double> h, s;
            if (!read_complex(gh, nm.c_str(), h) || !read_complex(gs, nm.c_str(), s) || h.size() != s.size())
                return bad("/matrix_hamilton, /matrix_overlap");
            d->hs.push_back(std::move(h));
            d->hs.push_back(std::move(s));
        }
        d->info.num_hs = (int32_t)(d->hs.size() / 2);
    }
    // final per-atom data written by hsgen_dump_write
    if (H5Lexists(f, "/hsgen", H5P_DEFAULT) > 0) {
        const int na = d->info.num_atoms, nl = d->info.max_lmax + 1;
        if (!read_doubles(f, "/hsgen/match", d->f_match) || d->f_match.size() != (size_t)na * nl * 5)
            return bad("/hsgen/match");
        if (!read_doubles(f, "/hsgen/rotmat", d->f_rot) || d->f_rot.size() != (size_t)na * 9)
            return bad("/hsgen/rotmat");
        H g(H5Gopen2(f, "/hsgen/tmat", H5P_DEFAULT), H5Gclose);
        if (!g.ok()) return bad("/hsgen/tmat");
        d->f_t.resize(na);
        for (int a = 0; a < na; ++a) {
            const std::string nm = std::to_string(a + 1);
            int64_t nel = 0;
            if (!read_complex(g, nm.c_str(), d->f_t[a], &nel)) return bad("/hsgen/tmat/<atom>");
            const int lm = d->lmax_t[d->atoms[a].type - 1];
            const int64_t n = (int64_t)(lm + 1) * (lm + 1);
            if (nel != 3 * n * n) return bad("/hsgen/tmat/<atom> size");
        }
        d->info.has_final = 1;
    }
    *out = d;
    return HSGEN_OK;
}

hsgen_status hsgen_dump_close(hsgen_dump* d)
{
    delete d;
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_info(const hsgen_dump* d, hsgen_dump_info* info)
{
    if (!d || !info) return fail(HSGEN_ERR_ARG, "null argument");
    *info = d->info;
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_type(const hsgen_dump* d, int32_t t, int32_t* lmax, int32_t* lnonsph,
                                 int32_t* mesh_n, double* dx, double* rmt, double* E, double* vr, double* t4)
{
    if (!d) return fail(HSGEN_ERR_ARG, "null dump");
    if (t < 0 || t >= d->info.num_types) return fail(HSGEN_ERR_ARG, "type %d out of range", t);
    const int lm = d->lmax_t[t];
    if (lmax) *lmax = lm;
    if (lnonsph) *lnonsph = d->lnonsph_t[t];
    if (mesh_n) *mesh_n = d->mesh[t].num;
    if (dx) *dx = d->mesh[t].dx;
    if (rmt) *rmt = d->mesh[t].rmt;
    if (E) std::memcpy(E, &d->energies[(size_t)t * (d->info.max_lmax + 1)], sizeof(double) * (lm + 1));
    if (vr)
        std::memcpy(vr, &d->potential[(size_t)t * d->info.max_mesh_pts], sizeof(double) * d->mesh[t].num);
    if (t4) std::memcpy(t4, d->tmat[t].data(), sizeof(double) * d->tmat[t].size());
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_atoms(const hsgen_dump* d, int32_t* type, double* pos, double* rot)
{
    if (!d) return fail(HSGEN_ERR_ARG, "null dump");
    for (int a = 0; a < d->info.num_atoms; ++a) {
        const AtomRec& r = d->atoms[a];
        if (type) type[a] = r.type - 1;
        if (pos)
            for (int i = 0; i < 3; ++i) pos[3 * a + i] = r.pos[i];
        if (rot)
            for (int i = 0; i < 9; ++i) rot[9 * a + i] = (double)r.rot[i];
    }
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_kpoint(const hsgen_dump* d, int32_t k, double* kpt, int32_t* ng, int32_t* g)
{
    if (!d) return fail(HSGEN_ERR_ARG, "null dump");
    if (k < 0 || k >= d->info.num_kpts) return fail(HSGEN_ERR_ARG, "k-point %d out of range", k);
    if (kpt)
        for (int i = 0; i < 3; ++i) kpt[i] = d->kpts[3 * (size_t)k + i];
    if (ng) *ng = d->num_gpts[k];
    if (g)
        std::memcpy(g, &d->gmesh[3 * (size_t)d->info.max_gpts * k], sizeof(int32_t) * 3 * (size_t)d->num_gpts[k]);
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_system(hsgen_dump* d, hsgen_system* desc)
{
    if (!d || !desc) return fail(HSGEN_ERR_ARG, "null argument");
    const int na = d->info.num_atoms, nl = d->info.max_lmax + 1;
    if (!d->built) {
        d->s_lmax.assign(na, 0);
        d->s_rmt.assign(na, 0.0);
        d->s_pos.assign(3 * (size_t)na, 0.0);
        d->s_rot.assign(9 * (size_t)na, 0.0);
        d->s_match.assign((size_t)na * nl * 5, 0.0);
        d->s_taa.assign(na, nullptr);
        d->s_tab.assign(na, nullptr);
        d->s_tbb.assign(na, nullptr);
        for (int a = 0; a < na; ++a) {
            const AtomRec& r = d->atoms[a];
            const int t = r.type - 1;
            d->s_lmax[a] = d->lmax_t[t];
            d->s_rmt[a] = d->mesh[t].rmt;
            for (int i = 0; i < 3; ++i) d->s_pos[3 * a + i] = r.pos[i];
        }
        if (d->info.has_final) {
            d->s_rot = d->f_rot;
            d->s_match = d->f_match;
            for (int a = 0; a < na; ++a) {
                const int64_t n = (int64_t)(d->s_lmax[a] + 1) * (d->s_lmax[a] + 1);
                d->s_taa[a] = d->f_t[a].data();
                d->s_tab[a] = d->f_t[a].data() + 2 * n * n;
                d->s_tbb[a] = d->f_t[a].data() + 4 * n * n;
            }
        } else {
            // rotation: the file's integers times the reciprocal basis, both
            // taken column-major (FleurData.cpp:421-429)
            for (int a = 0; a < na; ++a) {
                double rot[9];
                for (int i = 0; i < 9; ++i) rot[i] = (double)d->atoms[a].rot[i];
                matmul3(rot, d->info.rec_basis, &d->s_rot[9 * (size_t)a]);
            }
            // radial solutions per type (equal for all atoms of a type)
            const int nt = d->info.num_types;
            std::vector<double> tmatch((size_t)nt * nl * 5, 0.0);
            for (int t = 0; t < nt; ++t) {
                const MeshRec& m = d->mesh[t];
                const double* vr = &d->potential[(size_t)t * d->info.max_mesh_pts];
                const double* E = &d->energies[(size_t)t * nl];
                for (int l = 0; l <= d->lmax_t[t]; ++l) {
                    const hsgen_status st = hsgen_radial_solve(l, E[l], m.dx, m.rmt, m.num, vr,
                                                               &tmatch[((size_t)t * nl + l) * 5], nullptr, nullptr);
                    if (st) return fail(st, "radial solve failed for type %d, l = %d", t + 1, l);
                }
            }
            for (int a = 0; a < na; ++a) {
                const int t = d->atoms[a].type - 1;
                std::memcpy(&d->s_match[(size_t)a * nl * 5], &tmatch[(size_t)t * nl * 5], sizeof(double) * nl * 5);
            }
            // final T per type
            d->s_t.assign(nt, {});
            for (int t = 0; t < nt; ++t) {
                const int lm = d->lmax_t[t];
                const int64_t n = (int64_t)(lm + 1) * (lm + 1);
                std::vector<double> ddn(lm + 1);
                for (int l = 0; l <= lm; ++l) ddn[l] = tmatch[((size_t)t * nl + l) * 5 + 4];
                d->s_t[t].assign(6 * (size_t)n * n, 0.0);
                double* base = d->s_t[t].data();
                const hsgen_status st = hsgen_tmat_assemble(lm, d->lnonsph_t[t], d->tmat[t].data(),
                                                            &d->energies[(size_t)t * nl], ddn.data(), base,
                                                            base + 2 * n * n, base + 4 * n * n);
                if (st) return fail(st, "T assembly failed for type %d (t_ud and t_du diagonals differ?)", t + 1);
            }
            for (int a = 0; a < na; ++a) {
                const int t = d->atoms[a].type - 1;
                const int64_t n = (int64_t)(d->lmax_t[t] + 1) * (d->lmax_t[t] + 1);
                d->s_taa[a] = d->s_t[t].data();
                d->s_tab[a] = d->s_t[t].data() + 2 * n * n;
                d->s_tbb[a] = d->s_t[t].data() + 4 * n * n;
            }
        }
        d->built = true;
    }
    desc->natoms = na;
    desc->max_lmax = d->info.max_lmax;
    desc->lmax = d->s_lmax.data();
    desc->rmt = d->s_rmt.data();
    desc->pos = d->s_pos.data();
    desc->rotmat = d->s_rot.data();
    desc->match = d->s_match.data();
    desc->omega = d->info.omega;
    desc->bmat = d->info.rec_basis;
    desc->t_aa = d->s_taa.data();
    desc->t_ab = d->s_tab.data();
    desc->t_bb = d->s_tbb.data();
    return HSGEN_OK;
}

hsgen_status hsgen_dump_get_hs_packed(const hsgen_dump* d, int32_t k, int64_t* n, double* H, double* S)
{
    if (!d || !n) return fail(HSGEN_ERR_ARG, "null argument");
    if (k < 0 || k >= d->info.num_hs) return fail(HSGEN_ERR_ARG, "no H/S for k-point %d", k);
    const std::vector<double>& h = d->hs[2 * (size_t)k];
    const std::vector<double>& s = d->hs[2 * (size_t)k + 1];
    const int64_t m = (int64_t)h.size() / 2;                  // n (n + 1) / 2
    const int64_t dim = (int64_t)std::llround((std::sqrt(8.0 * (double)m + 1.0) - 1.0) / 2.0);
    if (dim * (dim + 1) / 2 != m) return fail(HSGEN_ERR_ARG, "packed size %lld is not triangular", (long long)m);
    *n = dim;
    if (H) std::memcpy(H, h.data(), sizeof(double) * h.size());
    if (S) std::memcpy(S, s.data(), sizeof(double) * s.size());
    return HSGEN_OK;
}

hsgen_status hsgen_unpack_lower(int64_t n, const double* packed, double* M, int64_t ld)
{
    if (n < 0 || !packed || !M || ld < n) return fail(HSGEN_ERR_ARG, "bad arguments");
    int64_t q = 0;
    for (int64_t i = 0; i < n; ++i)           // row by row, columns 0..i
        for (int64_t j = 0; j <= i; ++j, ++q) {
            M[2 * (j * ld + i)] = packed[2 * q];
            M[2 * (j * ld + i) + 1] = packed[2 * q + 1];
        }
    return HSGEN_OK;
}

// ---------------------------------------------------------------------------
// writers
// ---------------------------------------------------------------------------
hsgen_status hsgen_store_open(const char* path, hsgen_store** out)
{
    if (!path || !out) return fail(HSGEN_ERR_ARG, "null argument");
    quiet();
    hsgen_store* st = new hsgen_store();
    st->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (st->file >= 0) {
        st->gh = H5Gcreate2(st->file, "/matrix_hamilton", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        st->gs = H5Gcreate2(st->file, "/matrix_overlap", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    }
    if (st->file < 0 || st->gh < 0 || st->gs < 0) {
        hsgen_store_close(st);
        return fail(HSGEN_ERR_ARG, "cannot create %s", path);
    }
    *out = st;
    return HSGEN_OK;
}

hsgen_status hsgen_store_add_hs(hsgen_store* st, int32_t k, int32_t n, const double* Hm, const double* Sm,
                                int64_t ld)
{
    if (!st || !Hm || !Sm || n < 0 || ld < n || k < 0) return fail(HSGEN_ERR_ARG, "bad arguments");
    const hsize_t dims[2] = {(hsize_t)n, (hsize_t)n};
    H sp(H5Screate_simple(2, dims, nullptr), H5Sclose);
    H ct(cx_type(), H5Tclose);
    const std::string nm = std::to_string(k + 1);
    std::vector<double> buf(2 * (size_t)n * n);
    for (int w = 0; w < 2; ++w) {
        const double* src = w == 0 ? Hm : Sm;
        for (int64_t j = 0; j < n; ++j)
            std::memcpy(&buf[2 * (size_t)j * n], src + 2 * j * ld, sizeof(double) * 2 * n);
        H ds(H5Dcreate2(w == 0 ? st->gh : st->gs, nm.c_str(), ct, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
             H5Dclose);
        if (!ds.ok() || H5Dwrite(ds, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
            return fail(HSGEN_ERR_ARG, "cannot write k-point %d", k + 1);
    }
    return HSGEN_OK;
}

hsgen_status hsgen_store_close(hsgen_store* st)
{
    if (!st) return HSGEN_OK;
    if (st->gh >= 0) H5Gclose(st->gh);
    if (st->gs >= 0) H5Gclose(st->gs);
    if (st->file >= 0) H5Fclose(st->file);
    delete st;
    return HSGEN_OK;
}

hsgen_status hsgen_dump_write(const char* path, const hsgen_system* sys, const int32_t* types, const double* amat,
                              int32_t nk, const hsgen_kpoint* kps)
{
    if (!path || !sys || !types || !amat || nk < 0 || (nk > 0 && !kps)) return fail(HSGEN_ERR_ARG, "null argument");
    const int na = sys->natoms, nl = sys->max_lmax + 1;
    int nt = 0;
    for (int a = 0; a < na; ++a) nt = std::max(nt, types[a] + 1);
    std::vector<int> first(nt, -1);
    for (int a = 0; a < na; ++a) {
        if (types[a] < 0) return fail(HSGEN_ERR_ARG, "negative type");
        if (first[types[a]] < 0) first[types[a]] = a;
    }
    for (int t = 0; t < nt; ++t)
        if (first[t] < 0) return fail(HSGEN_ERR_ARG, "type %d has no atom", t);
    quiet();
    H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(HSGEN_ERR_ARG, "cannot create %s", path);
    auto wr = [&](hid_t loc, const char* name, hid_t ftype, hid_t mtype, int rank, const hsize_t* dims,
                  const void* data) -> bool {
        H sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
        H ds(H5Dcreate2(loc, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        return ds.ok() && H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    };
    const hsize_t d33[2] = {3, 3};
    bool ok = wr(f, "/real_basis", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, d33, amat) &&
              wr(f, "/rec_basis", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, d33, sys->bmat);
    // atoms: rotation R = rotmat * bmat^-1 must be integral for the layout
    double binv[9];
    {
        const double* b = sys->bmat;
        const double det = det3(b);
        if (!(std::fabs(det) > 0.0)) return fail(HSGEN_ERR_ARG, "singular reciprocal basis");
        binv[0] = (b[4] * b[8] - b[7] * b[5]) / det;
        binv[3] = -(b[3] * b[8] - b[6] * b[5]) / det;
        binv[6] = (b[3] * b[7] - b[6] * b[4]) / det;
        binv[1] = -(b[1] * b[8] - b[7] * b[2]) / det;
        binv[4] = (b[0] * b[8] - b[6] * b[2]) / det;
        binv[7] = -(b[0] * b[7] - b[6] * b[1]) / det;
        binv[2] = (b[1] * b[5] - b[4] * b[2]) / det;
        binv[5] = -(b[0] * b[5] - b[3] * b[2]) / det;
        binv[8] = (b[0] * b[4] - b[3] * b[1]) / det;
    }
    {
        std::vector<AtomRec> recs(na);
        for (int a = 0; a < na; ++a) {
            recs[a].type = types[a] + 1;
            for (int i = 0; i < 3; ++i) recs[a].pos[i] = sys->pos[3 * a + i];
            double R[9];
            matmul3(sys->rotmat + 9 * a, binv, R);
            for (int i = 0; i < 9; ++i) {
                const double r = std::nearbyint(R[i]);
                if (std::fabs(R[i] - r) > 1e-9) return fail(HSGEN_ERR_ARG, "atom %d: rotation is not integral", a);
                recs[a].rot[i] = (int32_t)r;
            }
        }
        const hsize_t d3[1] = {3};
        H a3m(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, d3), H5Tclose);
        H a33m(H5Tarray_create2(H5T_NATIVE_INT32, 2, d33), H5Tclose);
        H mt(H5Tcreate(H5T_COMPOUND, sizeof(AtomRec)), H5Tclose);
        H5Tinsert(mt, "type", HOFFSET(AtomRec, type), H5T_NATIVE_INT32);
        H5Tinsert(mt, "position", HOFFSET(AtomRec, pos), a3m);
        H5Tinsert(mt, "rotation", HOFFSET(AtomRec, rot), a33m);
        H a3f(H5Tarray_create2(H5T_IEEE_F64LE, 1, d3), H5Tclose);
        H a33f(H5Tarray_create2(H5T_STD_I32LE, 2, d33), H5Tclose);
        H ft(H5Tcreate(H5T_COMPOUND, 4 + 24 + 36), H5Tclose);
        H5Tinsert(ft, "type", 0, H5T_STD_I32LE);
        H5Tinsert(ft, "position", 4, a3f);
        H5Tinsert(ft, "rotation", 28, a33f);
        H5Tcommit2(f, "type_atoms", ft, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        const hsize_t dn[1] = {(hsize_t)na};
        H sp(H5Screate_simple(1, dn, nullptr), H5Sclose);
        H ds(H5Dcreate2(f, "/atoms", ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        ok = ok && ds.ok() && H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, recs.data()) >= 0 &&
             write_attr_int(ds, "num_atoms", na) && write_attr_int(ds, "num_types", nt);
    }
    // per type: cutoffs, mesh (no potential: num_points 0), zero energies
    {
        std::vector<int32_t> lm(nt);
        std::vector<MeshRec> ms(nt);
        for (int t = 0; t < nt; ++t) {
            lm[t] = sys->lmax[first[t]];
            ms[t] = MeshRec{0, 0.0, sys->rmt[first[t]]};
        }
        const hsize_t dt[1] = {(hsize_t)nt};
        ok = ok && wr(f, "/lmax", H5T_STD_I32LE, H5T_NATIVE_INT32, 1, dt, lm.data()) &&
             wr(f, "/lmax_nonspherical", H5T_STD_I32LE, H5T_NATIVE_INT32, 1, dt, lm.data());
        H mt(H5Tcreate(H5T_COMPOUND, sizeof(MeshRec)), H5Tclose);
        H5Tinsert(mt, "num_points", HOFFSET(MeshRec, num), H5T_NATIVE_INT32);
        H5Tinsert(mt, "dx", HOFFSET(MeshRec, dx), H5T_NATIVE_DOUBLE);
        H5Tinsert(mt, "rmt", HOFFSET(MeshRec, rmt), H5T_NATIVE_DOUBLE);
        H ft(H5Tcreate(H5T_COMPOUND, 20), H5Tclose);
        H5Tinsert(ft, "num_points", 0, H5T_STD_I32LE);
        H5Tinsert(ft, "dx", 4, H5T_IEEE_F64LE);
        H5Tinsert(ft, "rmt", 12, H5T_IEEE_F64LE);
        H5Tcommit2(f, "type_mesh", ft, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        H sp(H5Screate_simple(1, dt, nullptr), H5Sclose);
        H ds(H5Dcreate2(f, "/mesh", ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        ok = ok && ds.ok() && H5Dwrite(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, ms.data()) >= 0 &&
             write_attr_int(ds, "max_mesh_pts", 1);
        std::vector<double> zeros((size_t)nt * nl, 0.0);
        const hsize_t de[2] = {(hsize_t)nt, (hsize_t)nl}, dp[2] = {(hsize_t)nt, 1};
        ok = ok && wr(f, "/energy_params", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, de, zeros.data()) &&
             wr(f, "/potential", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, dp, zeros.data());
    }
    // k-points
    {
        int mg = 0;
        for (int k = 0; k < nk; ++k) mg = std::max(mg, kps[k].ng);
        const size_t rec = 4 + 24 + 12 * (size_t)mg;
        std::vector<char> buf(rec * (size_t)std::max(nk, 1), 0);
        for (int k = 0; k < nk; ++k) {
            char* r = buf.data() + rec * k;
            std::memcpy(r, &kps[k].ng, 4);
            std::memcpy(r + 4, kps[k].k, 24);
            std::memcpy(r + 28, kps[k].g, 12 * (size_t)kps[k].ng);
        }
        const hsize_t d3[1] = {3}, dg[2] = {(hsize_t)std::max(mg, 1), 3};
        H a3(H5Tarray_create2(H5T_IEEE_F64LE, 1, d3), H5Tclose);
        H ag(H5Tarray_create2(H5T_STD_I32LE, 2, dg), H5Tclose);
        H ft(H5Tcreate(H5T_COMPOUND, 4 + 24 + 12 * (size_t)std::max(mg, 1)), H5Tclose);
        H5Tinsert(ft, "num_gpts", 0, H5T_STD_I32LE);
        H5Tinsert(ft, "k_point", 4, a3);
        H5Tinsert(ft, "g_mesh", 28, ag);
        H5Tcommit2(f, "type_kmeshes", ft, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        buf.resize(std::max(buf.size(), (size_t)(4 + 24 + 12 * std::max(mg, 1)) * std::max(nk, 1)), 0);
        const hsize_t dn[1] = {(hsize_t)nk};
        H sp(H5Screate_simple(1, dn, nullptr), H5Sclose);
        H ds(H5Dcreate2(f, "/k_meshes", ft, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        ok = ok && ds.ok() && (nk == 0 || H5Dwrite(ds, ft, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0) &&
             write_attr_int(ds, "num_kpts", nk) && write_attr_int(ds, "max_gpts", mg);
    }
    // /tmat/<type>: lower triangles of the first atom's final T, with the
    // spherical diagonal terms the reference's loader adds (1/2 on T_ab at
    // zero energies) taken off
    {
        H g(H5Gcreate2(f, "/tmat", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && g.ok();
        for (int t = 0; t < nt && ok; ++t) {
            const int a = first[t];
            const int64_t n = (int64_t)(sys->lmax[a] + 1) * (sys->lmax[a] + 1);
            std::vector<double> buf(8 * (size_t)n * n, 0.0);
            const double* src[3] = {sys->t_aa[a], sys->t_bb[a], sys->t_ab[a]};
            for (int s = 0; s < 4; ++s) {
                double* o = buf.data() + 2 * (size_t)s * n * n;
                for (int64_t j = 0; j < n; ++j)
                    for (int64_t i = j; i < n; ++i) {
                        double re, im;
                        if (s < 3) {
                            re = src[s][2 * (j * n + i)];
                            im = src[s][2 * (j * n + i) + 1];
                        } else {   // t_du = (T_ab^H) lower: conj of T_ab(j, i)
                            re = sys->t_ab[a][2 * (i * n + j)];
                            im = -sys->t_ab[a][2 * (i * n + j) + 1];
                        }
                        if (s >= 2 && i == j) {
                            re = sys->t_ab[a][2 * (j * n + i)] - 0.5;
                            im = sys->t_ab[a][2 * (j * n + i) + 1];
                        }
                        o[2 * (j * n + i)] = re;
                        o[2 * (j * n + i) + 1] = im;
                    }
            }
            const hsize_t dnn[2] = {(hsize_t)n, (hsize_t)n}, d4[1] = {4};
            H ct(cx_type(), H5Tclose);
            H at(H5Tarray_create2(ct, 2, dnn), H5Tclose);
            const std::string nm = std::to_string(t + 1);
            H sp(H5Screate_simple(1, d4, nullptr), H5Sclose);
            H ds(H5Dcreate2(g, nm.c_str(), at, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
            ok = ds.ok() && H5Dwrite(ds, at, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) >= 0 &&
                 write_attr_int(ds, "n", (int32_t)n);
        }
    }
    // /hsgen: final per-atom data
    {
        H g(H5Gcreate2(f, "/hsgen", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        const hsize_t dm[3] = {(hsize_t)na, (hsize_t)nl, 5}, dr[2] = {(hsize_t)na, 9};
        ok = ok && g.ok() && wr(g, "match", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 3, dm, sys->match) &&
             wr(g, "rotmat", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, dr, sys->rotmat);
        H gt(H5Gcreate2(g, "tmat", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        ok = ok && gt.ok();
        for (int a = 0; a < na && ok; ++a) {
            const int64_t n = (int64_t)(sys->lmax[a] + 1) * (sys->lmax[a] + 1);
            std::vector<double> buf(6 * (size_t)n * n);
            std::memcpy(buf.data(), sys->t_aa[a], sizeof(double) * 2 * n * n);
            std::memcpy(buf.data() + 2 * n * n, sys->t_ab[a], sizeof(double) * 2 * n * n);
            std::memcpy(buf.data() + 4 * n * n, sys->t_bb[a], sizeof(double) * 2 * n * n);
            const hsize_t d3n[3] = {3, (hsize_t)n, (hsize_t)n};
            H ct(cx_type(), H5Tclose);
            const std::string nm = std::to_string(a + 1);
            ok = wr(gt, nm.c_str(), ct, ct, 3, d3n, buf.data());
        }
    }
    if (!ok) return fail(HSGEN_ERR_ARG, "writing %s failed", path);
    return HSGEN_OK;
}

hsgen_status hsgen_dump_write_hs_packed(const char* path, int32_t k, int32_t n, const double* Hm, const double* Sm,
                                        int64_t ld)
{
    if (!path || !Hm || !Sm || k < 0 || n < 0 || ld < std::max(n, 1)) return fail(HSGEN_ERR_ARG, "bad arguments");
    quiet();
    H f(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(HSGEN_ERR_ARG, "cannot open %s for writing", path);
    const std::string nm = std::to_string(k + 1);
    const hsize_t dims[1] = {(hsize_t)n * (hsize_t)(n + 1) / 2};
    std::vector<double> buf(2 * (size_t)dims[0]);
    for (int w = 0; w < 2; ++w) {
        const char* grp = w == 0 ? "/matrix_hamilton" : "/matrix_overlap";
        H g(H5Lexists(f, grp, H5P_DEFAULT) > 0 ? H5Gopen2(f, grp, H5P_DEFAULT)
                                                : H5Gcreate2(f, grp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Gclose);
        if (!g.ok()) return fail(HSGEN_ERR_ARG, "cannot create %s in %s", grp, path);
        if (H5Lexists(g, nm.c_str(), H5P_DEFAULT) > 0)
            return fail(HSGEN_ERR_ARG, "%s/%s exists in %s", grp, nm.c_str(), path);
        // the lower triangle row by row: (0,0), (1,0), (1,1), (2,0), ...
        const double* src = w == 0 ? Hm : Sm;
        size_t q = 0;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t j = 0; j <= i; ++j, ++q) {
                buf[2 * q] = src[2 * (j * ld + i)];
                buf[2 * q + 1] = src[2 * (j * ld + i) + 1];
            }
        H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
        H ct(cx_type(), H5Tclose);
        H ds(H5Dcreate2(g, nm.c_str(), ct, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        if (!ds.ok() || H5Dwrite(ds, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0)
            return fail(HSGEN_ERR_ARG, "cannot write %s/%s in %s", grp, nm.c_str(), path);
    }
    return HSGEN_OK;
}

}  // extern "C"
