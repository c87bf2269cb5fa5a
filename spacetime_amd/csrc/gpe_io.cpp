// Host-side files of libgpe: the reference solvers' state files (.dat) and their checkpoint
// pair, the legacy VTK files their viewers read, and their BMP density pictures.  Everything is written from (or read into)
// a level of a context through the ABI state layout psi[c][z][y][x][v] (gpe_get_state /
// gpe_set_state); the formats are the reference's, byte for byte where its arithmetic is
// correctly rounded (sqrt, division), so files move between the two codes.
#include "gpe.h"
#include "lattice_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

const double kDensityThreshold = 1e-4;  // DENSITY_THRESHOLD, Spin1GpeBCC/utils.h:789

// A state in the ABI layout on the host, with its grid.
struct Level {
    gpe_config cfg{};
    int V = 0, C = 0, nx = 0, ny = 0, nz = 0;
    const double *psi = nullptr;  // [c][z][y][x][v] x {re, im}
    std::vector<double> own;      // storage when the state was fetched from a context

    size_t unknown(int x, int y, int z, int v) const { return (((size_t)z * ny + y) * nx + x) * V + v; }
    size_t comp_stride() const { return (size_t)nz * ny * nx * V; }
    double re(int c, size_t u) const { return psi[2 * (c * comp_stride() + u)]; }
    double im(int c, size_t u) const { return psi[2 * (c * comp_stride() + u) + 1]; }
};

int shape_of(const gpe_config &cfg, Level &L)
{
    L.cfg = cfg;
    L.V = gpe_lattice_values(cfg.lattice);
    L.C = cfg.ncomp;
    L.nx = cfg.nx;
    L.ny = cfg.ny;
    L.nz = cfg.nz;
    if (L.V <= 0 || (L.C != 1 && L.C != 3 && L.C != 5) || L.nx < 1 || L.ny < 1 || L.nz < 1) return GPE_EINVAL;
    return GPE_OK;
}

int fetch(gpe_ctx *ctx, int level, Level &L)
{
    gpe_config cfg;
    int rc = gpe_get_config(ctx, &cfg);
    if (!rc) rc = shape_of(cfg, L);
    if (rc) return rc;
    try {
        L.own.assign(2 * (size_t)L.C * L.comp_stride(), 0.0);
    } catch (const std::bad_alloc &) {
        return GPE_ENOMEM;
    }
    L.psi = L.own.data();
    return gpe_get_state(ctx, level, L.own.data());
}

// Buffered binary writer; ok() false after any failed write.
class Out {
public:
    explicit Out(const char *path) : f_(path ? std::fopen(path, "wb") : nullptr) { buf_.reserve(1 << 20); }
    ~Out() { close(); }
    bool ok() const { return f_ != nullptr && good_; }
    void text(const std::string &s) { bytes(s.data(), s.size()); }
    void bytes(const void *p, size_t n)
    {
        const char *c = (const char *)p;
        buf_.insert(buf_.end(), c, c + n);
        if (buf_.size() >= (1u << 20)) flush();
    }
    // big-endian float32, as the reference's swapEnd + write (utils.h)
    void be_float(float v)
    {
        unsigned char b[4];
        std::memcpy(b, &v, 4);
        const unsigned char r[4] = {b[3], b[2], b[1], b[0]};
        bytes(r, 4);
    }
    bool close()
    {
        if (!f_) return false;
        flush();
        good_ = (std::fclose(f_) == 0) && good_;
        f_ = nullptr;
        return good_;
    }

private:
    void flush()
    {
        if (f_ && !buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) good_ = false;
        buf_.clear();
    }
    FILE *f_;
    bool good_ = true;
    std::vector<char> buf_;
};

// The reference's toString(double): fixed notation with 18 decimals (utils.h:102-108).
std::string fixed18(double t)
{
    char b[512];
    std::snprintf(b, sizeof(b), "%.18f", t);
    return b;
}

// ---------------------------------------------------------------- padded block files --

// Padded block array index of block (i, j, k) of the interior (halo = one block).
size_t padded_block(const Level &L, int i, int j, int k)
{
    return ((size_t)(k + 1) * (L.ny + 2) + (j + 1)) * (L.nx + 2) + (i + 1);
}

int write_dat(const Level &L, const char *path)
{
    const size_t nblocks = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    const size_t per_block = (size_t)L.V * L.C * 2;  // doubles
    std::vector<double> blocks;
    try {
        blocks.assign(nblocks * per_block, 0.0);
    } catch (const std::bad_alloc &) {
        return GPE_ENOMEM;
    }
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x)
                for (int v = 0; v < L.V; ++v) {
                    double *dst = &blocks[padded_block(L, x, y, z) * per_block + (size_t)v * L.C * 2];
                    const size_t u = L.unknown(x, y, z, v);
                    for (int c = 0; c < L.C; ++c) {
                        dst[2 * c] = L.re(c, u);
                        dst[2 * c + 1] = L.im(c, u);
                    }
                }
    Out o(path);
    if (!o.ok()) return GPE_EINVAL;
    o.bytes(blocks.data(), blocks.size() * sizeof(double));
    return o.close() ? GPE_OK : GPE_EDEVICE;
}

// 1 when the file exists and is readable.
bool readable(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// Read a padded block file of L's grid into psi (ABI layout).
int read_dat(const Level &L, const char *path, double *psi)
{
    const size_t nblocks = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    const size_t per_block = (size_t)L.V * L.C * 2;
    std::vector<double> blocks;
    try {
        blocks.assign(nblocks * per_block, 0.0);
    } catch (const std::bad_alloc &) {
        return GPE_ENOMEM;
    }
    FILE *f = path ? std::fopen(path, "rb") : nullptr;
    if (!f) return GPE_EINVAL;
    const size_t got = std::fread(blocks.data(), sizeof(double), blocks.size(), f);
    std::fclose(f);
    if (got != blocks.size()) return GPE_EINVAL;  // shorter than this grid's padded array
    for (int z = 0; z < L.nz; ++z)
        for (int y = 0; y < L.ny; ++y)
            for (int x = 0; x < L.nx; ++x)
                for (int v = 0; v < L.V; ++v) {
                    const double *src = &blocks[padded_block(L, x, y, z) * per_block + (size_t)v * L.C * 2];
                    const size_t u = L.unknown(x, y, z, v);
                    for (int c = 0; c < L.C; ++c) {
                        psi[2 * (c * L.comp_stride() + u)] = src[2 * c];
                        psi[2 * (c * L.comp_stride() + u) + 1] = src[2 * c + 1];
                    }
                }
    return GPE_OK;
}

int read_dat_into(gpe_ctx *ctx, int level, const char *path, double t)
{
    gpe_config cfg;
    Level L;
    int rc = gpe_get_config(ctx, &cfg);
    if (!rc) rc = shape_of(cfg, L);
    if (rc) return rc;
    std::vector<double> psi;
    try {
        psi.assign(2 * (size_t)L.C * L.comp_stride(), 0.0);
    } catch (const std::bad_alloc &) {
        return GPE_ENOMEM;
    }
    rc = read_dat(L, path, psi.data());
    return rc ? rc : gpe_set_state(ctx, level, psi.data(), t);
}

// ------------------------------------------------------------------------ VTK files --

// Points and values walk the padded grid in the reference's order: z, x, y, unknown.
template <typename F>
void walk_padded(const Level &L, F &&f)
{
    for (int z = 0; z < L.nz + 2; ++z)
        for (int x = 0; x < L.nx + 2; ++x)
            for (int y = 0; y < L.ny + 2; ++y)
                for (int v = 0; v < L.V; ++v) f(x, y, z, v);
}

bool interior(const Level &L, int x, int y, int z)
{
    return x > 0 && y > 0 && z > 0 && x < L.nx + 1 && y < L.ny + 1 && z < L.nz + 1;
}

// Spinor of padded point (x, y, z, v): zero on the halo; m = F..-F.
void spinor_at(const Level &L, int x, int y, int z, int v, double *re, double *im)
{
    if (!interior(L, x, y, z)) {
        for (int c = 0; c < L.C; ++c) re[c] = im[c] = 0.0;
        return;
    }
    const size_t u = L.unknown(x - 1, y - 1, z - 1, v);
    for (int c = 0; c < L.C; ++c) {
        re[c] = L.re(c, u);
        im[c] = L.im(c, u);
    }
}

// Spin-2 components in the quantisation basis of the reference's BASIS switch
// (Spin2GpeBCC/utils.h:1119-1147): Y and X quantised bases by the spin-2 rotation matrices.
void spin2_basis(int basis, double *re, double *im)
{
    if (basis == GPE_BASIS_Z) return;
    const double c = std::sqrt(6.0) * 0.25;
    const double r2 = re[0], r1 = re[1], r0 = re[2], rm1 = re[3], rm2 = re[4];
    const double i2 = im[0], i1 = im[1], i0 = im[2], im1 = im[3], im2 = im[4];
    if (basis == GPE_BASIS_X) {
        re[0] = 0.25 * r2 + 0.5 * r1 + c * r0 + 0.5 * rm1 + 0.25 * rm2;
        im[0] = 0.25 * i2 + 0.5 * i1 + c * i0 + 0.5 * im1 + 0.25 * im2;
        re[1] = -0.5 * r2 - 0.5 * r1 + 0.5 * rm1 + 0.5 * rm2;
        im[1] = -0.5 * i2 - 0.5 * i1 + 0.5 * im1 + 0.5 * im2;
        re[2] = c * r2 - 0.5 * r0 + c * rm2;
        im[2] = c * i2 - 0.5 * i0 + c * im2;
        re[3] = -0.5 * r2 + 0.5 * r1 - 0.5 * rm1 + 0.5 * rm2;
        im[3] = -0.5 * i2 + 0.5 * i1 - 0.5 * im1 + 0.5 * im2;
        re[4] = 0.25 * r2 - 0.5 * r1 + c * r0 - 0.5 * rm1 + 0.25 * rm2;
        im[4] = 0.25 * i2 - 0.5 * i1 + c * i0 - 0.5 * im1 + 0.25 * im2;
    } else {  // Y: i * (a + i b) = -b + i a
        re[0] = 0.25 * r2 + 0.5 * i1 - c * r0 - 0.5 * im1 + 0.25 * rm2;
        im[0] = 0.25 * i2 - 0.5 * r1 - c * i0 + 0.5 * rm1 + 0.25 * im2;
        re[1] = 0.5 * i2 - 0.5 * r1 - 0.5 * rm1 - 0.5 * im2;
        im[1] = -0.5 * r2 - 0.5 * i1 - 0.5 * im1 + 0.5 * rm2;
        re[2] = -c * r2 - 0.5 * r0 - c * rm2;
        im[2] = -c * i2 - 0.5 * i0 - c * im2;
        re[3] = -0.5 * i2 - 0.5 * r1 - 0.5 * rm1 + 0.5 * im2;
        im[3] = 0.5 * r2 - 0.5 * i1 - 0.5 * im1 - 0.5 * rm2;
        re[4] = 0.25 * r2 - 0.5 * i1 - c * r0 + 0.5 * im1 + 0.25 * rm2;
        im[4] = 0.25 * i2 + 0.5 * r1 - c * i0 - 0.5 * rm1 + 0.25 * im2;
    }
}

// Density summed over m = F..-F in the reference's order.
double density(const double *re, const double *im, int C)
{
    double n = 0.0;
    for (int c = 0; c < C; ++c) n += re[c] * re[c] + im[c] * im[c];
    return n;
}

void vtk_header(Out &o, uint64_t npoints)
{
    o.text("# vtk DataFile Version 3.0\nComment if needed\nBINARY\nDATASET POLYDATA\nPOINTS " +
           std::to_string(npoints) + " float\n");
}

void vtk_point(Out &o, const Level &L, int x, int y, int z, int v)
{
    const gpe::LatticeTable &lat = gpe::kLattices[L.cfg.lattice];
    const double s = L.cfg.block_scale;
    const int idx[3] = {x, y, z};
    for (int a = 0; a < 3; ++a)
        o.be_float((float)(L.cfg.p0[a] + s * (idx[a] * lat.block_width[a] + lat.pos[3 * v + a])));
}

// Label of component c: "1", "0", "-1" (spin-1), "2" ... "-2" (spin-2).
std::string m_label(int C, int c) { return std::to_string((C - 1) / 2 - c); }

// Spin-1 local spin vector of the reference's localAvgSpinAndDensity (Spin1GpeBCC.cu:235-271):
// the transverse part from the normalised spinor, the z part |psi_1|^2 - |psi_-1|^2 of the
// spinor itself.
void ref_local_spin(const double *re, const double *im, double out[3])
{
    const double n1 = re[0] * re[0] + im[0] * im[0];
    const double n0 = re[1] * re[1] + im[1] * im[1];
    const double nm = re[2] * re[2] + im[2] * im[2];
    const double dens = n1 + n0 + nm;
    const double sq = std::sqrt(dens);
    double r[3], i[3];
    for (int c = 0; c < 3; ++c) {
        r[c] = re[c] / sq;
        i[c] = im[c] / sq;
    }
    // sqrt2 (conj(p1) p0 + conj(p0) p-1)
    const double tr = (r[0] * r[1] + i[0] * i[1]) + (r[1] * r[2] + i[1] * i[2]);
    const double ti = (r[0] * i[1] - i[0] * r[1]) + (r[1] * i[2] - i[1] * r[2]);
    const double s2 = std::sqrt(2.0);
    out[0] = s2 * tr;
    out[1] = s2 * ti;
    out[2] = n1 - nm;
}

// Spin-1 u vector and theta of the reference's uvTheta (Spin1GpeBCC.cu:273-326).
void ref_u_theta(const double *re, const double *im, double u_out[3], double *theta_out)
{
    const double s2 = std::sqrt(2.0);
    // a = m + i n: ax = (p-1 - p1) / sqrt2, ay = -i (p-1 + p1) / sqrt2, az = p0
    const double axr = (re[2] - re[0]) / s2, axi = (im[2] - im[0]) / s2;
    const double ayr = (im[2] + im[0]) / s2, ayi = -(re[2] + re[0]) / s2;
    const double m[3] = {axr, ayr, re[1]}, n[3] = {axi, ayi, im[1]};
    const double mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
    const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double th = std::atan2(-2 * mn, mm - nn) / 2;
    if (th < 0) th += 3.14159265358979323846;
    const double st = std::sin(th), ct = std::cos(th);
    double u[3], w[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = m[k] * ct - st * n[k];
        w[k] = m[k] * st + ct * n[k];
    }
    const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double wn = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int k = 0; k < 3; ++k) u_out[k] = un >= wn ? u[k] : w[k];
    *theta_out = th;
}

int write_spinor(const Level &L, const char *path, int basis, bool preimage, int axis, double threshold)
{
    if (L.C < 3) return GPE_EINVAL;
    if (basis < 0 || basis > 2 || (L.C == 3 && basis != GPE_BASIS_Z)) return GPE_EINVAL;
    if (preimage && (L.C != 3 || axis < 0 || axis > 2)) return GPE_EINVAL;
    double re[5], im[5];
    // pre-image selection: interior points whose normalised local spin has |component axis|
    // above the threshold (savePreImageSpinor, Spin1GpeBCC/utils.h:1372-1390)
    auto selected = [&](int x, int y, int z, int v) -> bool {
        if (!preimage) return true;
        if (!interior(L, x, y, z)) return false;
        spinor_at(L, x, y, z, v, re, im);
        const double n1 = re[0] * re[0] + im[0] * im[0], nm = re[2] * re[2] + im[2] * im[2];
        const double s2 = std::sqrt(2.0);
        const double tx = s2 * ((re[0] * re[1] + im[0] * im[1]) + (re[1] * re[2] + im[1] * im[2]));
        const double ty = s2 * ((re[0] * im[1] - im[0] * re[1]) + (re[1] * im[2] - im[1] * re[2]));
        const double sp[3] = {tx, ty, n1 - nm};
        const double len = std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]);
        return std::fabs(sp[axis] / len) > threshold;  // 0 / 0 (no spin) is not selected
    };
    uint64_t npoints = 0;
    walk_padded(L, [&](int x, int y, int z, int v) { npoints += selected(x, y, z, v) ? 1 : 0; });
    Out o(path);
    if (!o.ok()) return GPE_EINVAL;
    vtk_header(o, npoints);
    walk_padded(L, [&](int x, int y, int z, int v) {
        if (selected(x, y, z, v)) vtk_point(o, L, x, y, z, v);
    });
    // spin-1 files have a blank line after POINT_DATA, spin-2 files none
    o.text("\nPOINT_DATA " + std::to_string(npoints) + "\n" + (L.C == 3 ? "\n" : ""));
    for (int c = 0; c < L.C; ++c)
        for (int part = 0; part < 2; ++part) {
            if (c > 0 || part > 0) o.text("\n");
            o.text(std::string("SCALARS ") + (part ? "i" : "r") + "_m=" + m_label(L.C, c) + " float 1\nLOOKUP_TABLE default\n");
            walk_padded(L, [&](int x, int y, int z, int v) {
                if (!selected(x, y, z, v)) return;
                spinor_at(L, x, y, z, v, re, im);
                if (L.C == 5) spin2_basis(basis, re, im);
                const double dens = density(re, im, L.C);
                float val = 0.f;
                if (kDensityThreshold < dens) val = (float)((part ? im[c] : re[c]) / std::sqrt(dens));
                o.be_float(val);
            });
        }
    o.text("\n");
    return o.close() ? GPE_OK : GPE_EDEVICE;
}

int write_volume(const Level &L, const char *path, int basis)
{
    if (L.C < 3 || basis < 0 || basis > 2 || (L.C == 3 && basis != GPE_BASIS_Z)) return GPE_EINVAL;
    const uint64_t npoints = (uint64_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2) * L.V;
    Out o(path);
    if (!o.ok()) return GPE_EINVAL;
    vtk_header(o, npoints);
    walk_padded(L, [&](int x, int y, int z, int v) { vtk_point(o, L, x, y, z, v); });
    o.text("\nPOINT_DATA " + std::to_string(npoints) + "\n");
    double re[5], im[5];
    if (L.C == 5) {
        // component densities |psi_m|^2 in the chosen basis (Spin2GpeBCC/utils.h:850-1066)
        for (int c = 0; c < 5; ++c) {
            o.text(std::string(c ? "\n" : "") + "SCALARS m=" + m_label(5, c) + " float 1\nLOOKUP_TABLE default\n");
            walk_padded(L, [&](int x, int y, int z, int v) {
                spinor_at(L, x, y, z, v, re, im);
                spin2_basis(basis, re, im);
                o.be_float((float)(re[c] * re[c] + im[c] * im[c]));
            });
        }
        o.text("\n");
        return o.close() ? GPE_OK : GPE_EDEVICE;
    }
    // spin-1: density, spinNorm, theta, localAvgSpin, u (Spin1GpeBCC/utils.h:832-1072)
    o.text("SCALARS density float 1\nLOOKUP_TABLE default\n");
    walk_padded(L, [&](int x, int y, int z, int v) {
        spinor_at(L, x, y, z, v, re, im);
        o.be_float((float)density(re, im, 3));
    });
    const int spin_norm = 1, theta_f = 2, spin_vec = 3, u_vec = 4;
    auto emit = [&](int which) {
        walk_padded(L, [&](int x, int y, int z, int v) {
            spinor_at(L, x, y, z, v, re, im);
            const double dens = density(re, im, 3);
            const bool live = dens > kDensityThreshold && interior(L, x, y, z);
            double sp[3] = {0, 0, 0}, u[3] = {0, 0, 0}, th = 0.0;
            if (live) {
                ref_local_spin(re, im, sp);
                ref_u_theta(re, im, u, &th);
            }
            if (which == spin_norm) {
                o.be_float(live ? (float)std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]) : 0.f);
            } else if (which == theta_f) {
                o.be_float((float)th);
            } else {
                const double *w = which == spin_vec ? sp : u;
                for (int k = 0; k < 3; ++k) o.be_float((float)w[k]);
            }
        });
    };
    o.text("\nSCALARS spinNorm float 1\nLOOKUP_TABLE default\n");
    emit(spin_norm);
    o.text("\nSCALARS theta float 1\nLOOKUP_TABLE default\n");
    emit(theta_f);
    o.text("\nVECTORS localAvgSpin float\n");
    emit(spin_vec);
    o.text("\nVECTORS u float\n");
    emit(u_vec);
    o.text("\n");
    return o.close() ? GPE_OK : GPE_EDEVICE;
}

// ---------------------------------------------------------------------- BMP pictures --

// Spin-1 components in the quantisation basis of the reference's BASIS switch
// (Spin1GpeBCC/utils.h:176-190): s' = U s with the rotation of the m = 1, 0, -1 basis onto the
// x or y axis, evaluated as the reference writes it (halves, then the sqrt 2 divisions;
// s0 * (-i) = (im, -re) exactly).
void spin1_basis(int basis, double *re, double *im)
{
    if (basis == GPE_BASIS_Z) return;
    const double s2 = std::sqrt(2.0);
    const double r1 = re[0], r0 = re[1], rm = re[2];
    const double i1 = im[0], i0 = im[1], imm = im[2];
    if (basis == GPE_BASIS_X) {
        re[0] = 0.5 * (r1 + rm) - r0 / s2;
        im[0] = 0.5 * (i1 + imm) - i0 / s2;
        re[1] = (r1 - rm) / s2;
        im[1] = (i1 - imm) / s2;
        re[2] = 0.5 * (rm + r1) + r0 / s2;
        im[2] = 0.5 * (imm + i1) + i0 / s2;
    } else {
        re[0] = 0.5 * (r1 - rm) + i0 / s2;
        im[0] = 0.5 * (i1 - imm) + (-r0) / s2;
        re[1] = (i1 + imm) / s2;
        im[1] = (-(r1 + rm)) / s2;
        re[2] = 0.5 * (rm - r1) + i0 / s2;
        im[2] = 0.5 * (imm - i1) + (-r0) / s2;
    }
}

// An RGB picture saved as the reference's Picture::save writes it (Output/Picture.cpp:83-138):
// "BM", a 40-byte info header, 24-bit pixels as blue, green, red bytes, rows from y = 0 up, each
// padded to 4 bytes; a channel value v is stored as 256 v truncated, 0 below 1 and 255 above
// 255.  The header's file size and image size count 4 bytes per pixel, as the reference's do,
// so the files match it byte for byte.  Pixels start black.
class Picture {
public:
    Picture(int w, int h) : w_(w), h_(h), px_((size_t)w * h * 3, 0.0) {}
    int width() const { return w_; }
    int height() const { return h_; }
    // Writes outside the picture are dropped (the reference's setColor drops x >= width).
    void set(int x, int y, double r, double g, double b)
    {
        if (x < 0 || y < 0 || x >= w_ || y >= h_) return;
        double *p = &px_[3 * ((size_t)y * w_ + x)];
        p[0] = r;
        p[1] = g;
        p[2] = b;
    }
    int save(const char *path) const
    {
        Out o(path);
        if (!o.ok()) return GPE_EINVAL;
        const uint32_t offs = 14 + 40, data = 4u * (uint32_t)((size_t)w_ * h_), size = offs + data;
        const uint32_t zero = 0, info = 40, w = (uint32_t)w_, h = (uint32_t)h_;
        const uint16_t planes = 1, bits = 24;
        o.bytes("BM", 2);
        o.bytes(&size, 4);
        o.bytes(&zero, 4);
        o.bytes(&offs, 4);
        o.bytes(&info, 4);
        o.bytes(&w, 4);
        o.bytes(&h, 4);
        o.bytes(&planes, 2);
        o.bytes(&bits, 2);
        o.bytes(&zero, 4);
        o.bytes(&data, 4);
        for (int k = 0; k < 4; ++k) o.bytes(&zero, 4);
        const int pad = 3 - ((3 * w_ + 3) % 4);
        std::vector<unsigned char> row((size_t)3 * w_ + pad, 0);
        for (int y = 0; y < h_; ++y) {
            const double *p = &px_[3 * (size_t)y * w_];
            for (int x = 0; x < w_; ++x) {
                row[3 * x + 0] = channel(p[3 * x + 2]);
                row[3 * x + 1] = channel(p[3 * x + 1]);
                row[3 * x + 2] = channel(p[3 * x + 0]);
            }
            o.bytes(row.data(), row.size());
        }
        return o.close() ? GPE_OK : GPE_EDEVICE;
    }

private:
    static unsigned char channel(double v)
    {
        const double c = 256.0 * v;
        if (!(c >= 1.0)) return 0;  // also NaN, which the reference's conversion turns into 0
        if (c > 255.0) return 255;
        return (unsigned char)(int)c;
    }
    int w_, h_;
    std::vector<double> px_;
};

// Densities of the reference's pictures projected onto a plane of the padded block grid: for
// block (a, b) of the plane's two axes and component m, the sums over the third axis of
// |psi_m|^2, of Re^2 and of Im^2 (in the picture's basis).  The reference sums pixel by pixel
// with two pixels per block edge, so every block of the column enters twice in a row, each
// time with its unknowns v = 0..V-1 in turn; the sums keep that order.  Halo blocks hold
// zeros, and skipping them leaves every sum unchanged.
struct Plane {
    int na = 0, nb = 0, C = 0;
    std::vector<double> n, r, i;  // [(b * na + a) * C + m]
    size_t at(int a, int b, int m) const { return ((size_t)b * na + a) * C + m; }
    // Largest projected density, folded as the reference's nested std::max (utils.h:199).
    double max_density() const
    {
        double mx = 0.0;
        for (size_t k = 0; k < n.size(); k += C) {
            double m = n[k + C - 1];
            for (int c = C - 2; c >= 0; --c) m = std::max(n[k + c], m);
            mx = std::max(mx, m);
        }
        return mx;
    }
};

// axis numbers 0 = x, 1 = y, 2 = z; a, b = the picture's horizontal and vertical axes.
Plane project(const Level &L, int a, int b, int basis)
{
    const int P[3] = {L.nx + 2, L.ny + 2, L.nz + 2};
    const int s = 3 - a - b;
    Plane pl;
    pl.na = P[a];
    pl.nb = P[b];
    pl.C = L.C;
    const size_t cells = (size_t)pl.na * pl.nb * L.C;
    pl.n.assign(cells, 0.0);
    pl.r.assign(cells, 0.0);
    pl.i.assign(cells, 0.0);
    std::vector<double> tn((size_t)L.V * L.C), tr((size_t)L.V * L.C), ti((size_t)L.V * L.C);
    double re[5], im[5];
    for (int qb = 1; qb < P[b] - 1; ++qb)
        for (int qa = 1; qa < P[a] - 1; ++qa) {
            double *n = &pl.n[pl.at(qa, qb, 0)], *r = &pl.r[pl.at(qa, qb, 0)], *i = &pl.i[pl.at(qa, qb, 0)];
            for (int qs = 1; qs < P[s] - 1; ++qs) {
                int q[3];
                q[a] = qa;
                q[b] = qb;
                q[s] = qs;
                for (int v = 0; v < L.V; ++v) {
                    spinor_at(L, q[0], q[1], q[2], v, re, im);
                    if (L.C == 3) spin1_basis(basis, re, im);
                    if (L.C == 5) spin2_basis(basis, re, im);
                    for (int c = 0; c < L.C; ++c) {
                        const double rr = re[c] * re[c], ii = im[c] * im[c];
                        tr[v * L.C + c] = rr;
                        ti[v * L.C + c] = ii;
                        tn[v * L.C + c] = rr + ii;
                    }
                }
                for (int rep = 0; rep < 2; ++rep)
                    for (int v = 0; v < L.V; ++v)
                        for (int c = 0; c < L.C; ++c) {
                            n[c] += tn[v * L.C + c];
                            r[c] += tr[v * L.C + c];
                            i[c] += ti[v * L.C + c];
                        }
            }
        }
    return pl;
}

// Panels of one plane: component m's panel starts at x = m * pw; pixel (u, w) of the panel
// shows block (u / 2, w / 2) of the plane at picture pixel (x0 + u', y0 + w') with u' = u or
// pw - 1 - u (flip_u) and w' likewise.  ri: red = I * sum Re^2, green = I * sum Im^2; else a
// grey level I * sum |psi_m|^2.
void draw_plane(Picture &pic, const Plane &pl, int pw, int y0, double I, bool ri, bool flip_u, bool flip_w)
{
    const int nu = 2 * pl.na, nw = 2 * pl.nb;
    for (int w = 0; w < nw; ++w)
        for (int u = 0; u < nu; ++u)
            for (int c = 0; c < pl.C; ++c) {
                const size_t k = pl.at(u / 2, w / 2, c);
                const int x = c * pw + (flip_u ? nu - 1 - u : u), y = y0 + (flip_w ? nw - 1 - w : w);
                if (ri)
                    pic.set(x, y, I * pl.r[k], I * pl.i[k], 0.0);
                else {
                    const double g = I * pl.n[k];
                    pic.set(x, y, g, g, g);
                }
            }
}

// Grey separators: rows at y = rows[k], columns at x = k * pw (k = 1..panels-1).
void separators(Picture &pic, int pw, int panels, const int *rows, int nrows)
{
    for (int k = 0; k < nrows; ++k)
        for (int x = 0; x < pic.width(); ++x) pic.set(x, rows[k], 0.5, 0.5, 0.5);
    for (int y = 0; y < pic.height(); ++y)
        for (int k = 1; k < panels; ++k) pic.set(k * pw, y, 0.5, 0.5, 0.5);
}

// The reference's density pictures of a state, with W = 2 (nx + 2), H = 2 (ny + 2),
// D = 2 (nz + 2) pixels along the padded grid:
//   GPE_PICTURE_DENSITY, spin-1 (drawDensity, Spin1GpeBCC/utils.h:151-366): 3 W x 2 H; top
//     row of panels (y < D) the column sums over y (XZ), bottom row (y >= H) the sums over z
//     (XY), both mirrored in x and the XY panels in y; grey |psi_m|^2 / max of the plane.
//   GPE_PICTURE_DENSITY, spin-2 (drawDensity, Spin2GpeBCC/utils.h:287-796): 5 W x 3 H; XZ,
//     then YZ (sums over x, scaled by the XZ maximum as the reference does), then XY, unmirrored.
//   GPE_PICTURE_RI, spin-1 (drawDensityRI, Spin1GpeBCC/utils.h:368-580): 3 W x 2 H, XZ and XY
//     unmirrored, red = sum Re^2, green = sum Im^2, both times 2 / the plane's largest sum of
//     |psi_m|^2.
//   GPE_PICTURE_RI, spin-2 (drawIandR, Spin2GpeBCC/utils.h:131-285): 5 W x 2 H, XZ and XY,
//     red = sum Re^2, green = sum Im^2 unscaled, components in the z basis.
//   GPE_PICTURE_SLICE, scalar (GPUGrossPitaevskii_BCC.cu:700-723): W x H, the padded block
//     layer z = nz / 2 + 1, unknown v = 0: red = 20 Re^2, green = 20 Im^2.
int write_picture(const Level &L, const char *path, int style, int basis)
{
    if (basis < 0 || basis > 2) return GPE_EINVAL;
    const int W = 2 * (L.nx + 2), H = 2 * (L.ny + 2);
    if (style == GPE_PICTURE_SLICE) {
        if (L.C != 1) return GPE_EINVAL;
        Picture pic(W, H);
        const int k = L.nz / 2 + 1;
        for (int j = 0; j < H; ++j)
            for (int i = 0; i < W; ++i) {
                double re[1], im[1];
                spinor_at(L, i / 2, j / 2, k, 0, re, im);
                pic.set(i, j, 20.0 * (re[0] * re[0]), 20.0 * (im[0] * im[0]), 0.0);
            }
        return pic.save(path);
    }
    if ((style != GPE_PICTURE_DENSITY && style != GPE_PICTURE_RI) || L.C == 1) return GPE_EINVAL;
    const bool ri = style == GPE_PICTURE_RI;
    if (ri && L.C == 5 && basis != GPE_BASIS_Z) return GPE_EINVAL;
    const Plane xz = project(L, 0, 2, basis);
    const Plane xy = project(L, 0, 1, basis);
    if (L.C == 3) {
        Picture pic(3 * W, 2 * H);
        const double num = ri ? 2.0 : 1.0;
        draw_plane(pic, xz, W, 0, num / xz.max_density(), ri, !ri, false);
        draw_plane(pic, xy, W, H, num / xy.max_density(), ri, !ri, !ri);
        const int rows[1] = {H};
        separators(pic, W, 3, rows, 1);
        return pic.save(path);
    }
    if (ri) {
        Picture pic(5 * W, 2 * H);
        draw_plane(pic, xz, W, 0, 1.0, true, false, false);
        draw_plane(pic, xy, W, H, 1.0, true, false, false);
        const int rows[1] = {H};
        separators(pic, W, 5, rows, 1);
        return pic.save(path);
    }
    Picture pic(5 * W, 3 * H);
    const double Ixz = 1.0 / xz.max_density();
    draw_plane(pic, xz, W, 0, Ixz, false, false, false);
    draw_plane(pic, project(L, 1, 2, basis), W, H, Ixz, false, false, false);
    draw_plane(pic, xy, W, 2 * H, 1.0 / xy.max_density(), false, false, false);
    const int rows[2] = {H, 2 * H};
    separators(pic, W, 5, rows, 2);
    return pic.save(path);
}

std::string join(const char *dir, const char *prefix, const std::string &name)
{
    std::string p = dir && *dir ? std::string(dir) + "/" : std::string();
    return p + (prefix ? prefix : "") + name;
}

}  // namespace

extern "C" {

int gpe_host_write_dat(const gpe_config *cfg, const double *psi, const char *path)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    L.psi = psi;
    return write_dat(L, path);
}

int gpe_host_read_dat(const gpe_config *cfg, const char *path, double *psi)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    return read_dat(L, path, psi);
}

int gpe_host_write_vtk_spinor(const gpe_config *cfg, const double *psi, const char *path, int basis)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    L.psi = psi;
    return write_spinor(L, path, basis, false, 0, 0.0);
}

int gpe_host_write_vtk_preimage(const gpe_config *cfg, const double *psi, const char *path, int axis,
                                double threshold)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    L.psi = psi;
    return write_spinor(L, path, GPE_BASIS_Z, true, axis, threshold);
}

int gpe_host_write_vtk_volume(const gpe_config *cfg, const double *psi, const char *path, int basis)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    L.psi = psi;
    return write_volume(L, path, basis);
}

int gpe_write_dat(gpe_ctx *ctx, int level, const char *path)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    Level L;
    int rc = fetch(ctx, level, L);
    return rc ? rc : write_dat(L, path);
}

int gpe_read_dat(gpe_ctx *ctx, int level, const char *path, double t)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    return read_dat_into(ctx, level, path, t);
}

int gpe_save_checkpoint(gpe_ctx *ctx, const char *dir, const char *prefix, char *tag, int tag_cap)
{
    if (!ctx) return GPE_EINVAL;
    double t = 0.0, t0 = 0.0;
    long long sweeps = 0;
    int rc = gpe_time(ctx, &t, &sweeps);
    if (!rc) rc = gpe_clock(ctx, &t0, &sweeps);
    if (rc) return rc;
    const std::string T = fixed18(t);
    Level L;
    rc = fetch(ctx, 1, L);
    if (!rc) rc = write_dat(L, join(dir, prefix, T + ".dat").c_str());
    if (!rc) rc = fetch(ctx, 0, L);
    if (!rc) rc = write_dat(L, join(dir, prefix, "even_" + T + ".dat").c_str());
    if (rc) return rc;
    FILE *f = std::fopen(join(dir, prefix, T + ".clock").c_str(), "w");
    if (!f) return GPE_EINVAL;
    const bool ok = std::fprintf(f, "%a %lld\n", t0, sweeps) > 0;
    if (std::fclose(f) != 0 || !ok) return GPE_EDEVICE;
    if (tag && tag_cap > 0) {
        std::strncpy(tag, T.c_str(), (size_t)tag_cap - 1);
        tag[tag_cap - 1] = '\0';
    }
    return GPE_OK;
}

int gpe_load_checkpoint(gpe_ctx *ctx, const char *dir, const char *prefix, const char *tag, int *levels)
{
    if (!ctx || !tag) return GPE_EINVAL;
    const double t = std::strtod(tag, nullptr);
    const std::string odd = join(dir, prefix, std::string(tag) + ".dat");
    const std::string even = join(dir, prefix, std::string("even_") + tag + ".dat");
    if (!readable(odd)) return GPE_EINVAL;
    int rc, nlev;
    if (readable(even)) {
        rc = read_dat_into(ctx, 0, even.c_str(), t);
        if (!rc) rc = read_dat_into(ctx, 1, odd.c_str(), t);
        nlev = 2;
    } else {
        rc = read_dat_into(ctx, 0, odd.c_str(), t);
        nlev = 1;
    }
    if (rc) return rc;
    FILE *f = std::fopen(join(dir, prefix, std::string(tag) + ".clock").c_str(), "r");
    if (f) {
        double t0 = 0.0;
        long long sweeps = 0;
        const int n = std::fscanf(f, "%la %lld", &t0, &sweeps);
        std::fclose(f);
        if (n != 2) return GPE_EINVAL;
        rc = gpe_set_time(ctx, t0, sweeps);
        if (rc) return rc;
    }
    if (levels) *levels = nlev;
    return GPE_OK;
}

int gpe_write_vtk_spinor(gpe_ctx *ctx, int level, const char *path, int basis)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    Level L;
    int rc = fetch(ctx, level, L);
    return rc ? rc : write_spinor(L, path, basis, false, 0, 0.0);
}

int gpe_write_vtk_preimage(gpe_ctx *ctx, int level, const char *path, int axis, double threshold)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    Level L;
    int rc = fetch(ctx, level, L);
    return rc ? rc : write_spinor(L, path, GPE_BASIS_Z, true, axis, threshold);
}

int gpe_write_vtk_volume(gpe_ctx *ctx, int level, const char *path, int basis)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    Level L;
    int rc = fetch(ctx, level, L);
    return rc ? rc : write_volume(L, path, basis);
}

int gpe_write_bmp(gpe_ctx *ctx, int level, const char *path, int style, int basis)
{
    if (!ctx || !path || (level != 0 && level != 1)) return GPE_EINVAL;
    Level L;
    int rc = fetch(ctx, level, L);
    return rc ? rc : write_picture(L, path, style, basis);
}

int gpe_host_write_bmp(const gpe_config *cfg, const double *psi, const char *path, int style, int basis)
{
    Level L;
    if (!cfg || !psi || !path || shape_of(*cfg, L)) return GPE_EINVAL;
    L.psi = psi;
    return write_picture(L, path, style, basis);
}

}  // extern "C"
