// dg_cli -- the reference's command-line host (src/rs/src/main.rs) on the
// MI355X step library: the same demo set-ups (src/rs/src/init.rs, with their
// ghost-inclusive coordinates), the same loop and output cadence
// (main.rs:38-68: dump after the step when wtime > tout or on the first
// step, then tnow += dt in f32), the state written per dump.
//
//   dg_cli --demo KELVIN_HELMHOLTZ --tout 0.1 --total-time 1.0 --max-steps 100
//          [--precision fp32|fp64|fp32_bulk] [--math exact|fast]
//          [--limiter none|minmod|vanleer] [--out DIR] [--no-dump] [--quiet]
//          [--restart DIR/state.NNNNN]
//   dg_cli --config sod64|blast256|khi512|ensemble|learned128 --steps N [--seed S]
//
// Dumps: DIR/state.NNNNN/ (main.rs:58-60) with one .npy file per dataset of
// the reference's HDF5 layout (io.rs:5-196), float64, full ghosted shape:
// primitives.{rho,vx,vy,vz,pressure} (those of the start of the dumped step,
// as the reference grid holds them after compute_step), conserved.{mass,
// momentum_x,momentum_y,momentum_z,energy} (the advanced state; ghost cells
// follow the boundary maps), gradients.{rho,vx,vy,vz,pressure} and
// fluxes.{mass,momentum_x,momentum_y,momentum_z,energy} of the dumped step,
// shape (nx+4, ny+4, nz+4, 3) with the axis last.  libhdf5 is not available;
// .npy is read by numpy and every HDF5-side tool chain through it.  Each dump
// also holds meta.txt: the loop clock after the dumped step (step, tnow,
// wtime, wstep), from which --restart continues the run exactly.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/deepgreen.h"

namespace {

// reference constants (globals.rs)
const float kG = -0.1f, kDensity = 1.293f, kPressure = 101325.0f, kRs = 287.0f;
const float kPi = 3.14159265358979323846f;

struct Grid {
    int nx, ny, nz;  // ghosted sizes (Simulation::nx etc., grid.rs:64-66)
    float ds;
    std::vector<float> f[5];  // rho, vx, vy, vz, p
    Grid(int ix, int iy, int iz, float d) : nx(ix + 4), ny(iy + 4), nz(iz + 4), ds(d)
    {
        for (auto &v : f) v.assign((size_t)nx * ny * nz, 0.0f);
    }
    float &at(int q, int i, int j, int k) { return f[q][((size_t)i * ny + j) * nz + k]; }
    void fill(int q, float v) { for (auto &x : f[q]) x = v; }
};

struct Demo {
    int n[3];
    float ds;
    int bc[6];
    bool gravity;
    std::function<void(Grid &)> init;
};

enum { P = DG_PERIODIC, O = DG_OUTFLOW, W = DG_WALL };

// init_khi (init.rs:401-444)
void init_khi(Grid &g)
{
    const float kappa = 0.1f, s = 0.05f / std::sqrt(2.0f);
    g.fill(4, 2.5f);
    g.fill(2, 0.0f);
    const float H = g.ds * g.nz;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float xs = i * g.ds, zs = k * g.ds;
                const float in = std::fabs(zs - H / 2.0f) < H / 4.0f ? 1.0f : 0.0f;
                g.at(0, i, j, k) = 1.0f + in;
                g.at(1, i, j, k) = -0.5f + in;
                g.at(3, i, j, k) = kappa * std::sin(4.0f * kPi * xs) *
                                   (std::exp(-std::pow(zs - H / 4.0f, 2.0f) / (2.0f * s * s)) +
                                    std::exp(-std::pow(zs - 3.0f * H / 4.0f, 2.0f) / (2.0f * s * s)));
            }
}

// the two lowest z planes of p from plane 2 plus hydrostatic increments
void hydrostatic_floor(Grid &g)
{
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j) {
            const float p2 = g.at(4, i, j, 2);
            g.at(4, i, j, 1) = p2 + std::fabs(kG) * g.ds;
            g.at(4, i, j, 0) = p2 + 2.0f * std::fabs(kG) * g.ds;
        }
}

// init_rt (init.rs:446-486)
void init_rt(Grid &g)
{
    const float xmin = -0.25f, zmin = -0.75f, p0 = 2.5f;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float xs = xmin + g.ds * i, zs = zmin + g.ds * k;
                const float rho = zs > 0.0f ? 2.0f : 1.0f;
                g.at(0, i, j, k) = rho;
                g.at(4, i, j, k) = p0 - 0.1f * rho * 1.4f;
                g.at(3, i, j, k) =
                    -0.01f * (1.0f + std::cos(4.0f * kPi * xs)) * (1.0f + std::cos(3.0f * kPi * zs)) / 4.0f;
            }
    hydrostatic_floor(g);
}

// init_rb (init.rs:323-366)
void init_rb(Grid &g)
{
    const float rho0 = 1.0f, rsp = 287.052874f, t0 = 373.0f, tf = t0 - 20.0f;
    const float p0 = rho0 * rsp * t0, pf = rho0 * rsp * tf;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j) {
            float noise = (std::cos(8.0f * kPi * j / g.ny) + std::sin(6.0f * kPi * i / g.nx)) * 0.05f;
            for (int k = 0; k < g.nz; ++k) {
                const float zs = k * g.ds;
                if (!(zs < 0.2f)) noise = 0.0f;
                g.at(0, i, j, k) = rho0 + noise;
                g.at(4, i, j, k) = p0 + ((pf - p0) / (g.nz * g.ds) * zs);
            }
        }
}

// init_explosion_2d (init.rs:696-724)
void init_explosion_2d(Grid &g)
{
    g.fill(0, 1.0f);
    g.fill(4, 2.5f);
    const float sigma = 0.0001f;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float xs = ((i - g.nx / 2.0f) * g.ds) / g.nx;
                const float ys = ((j - g.ny / 2.0f) * g.ds) / g.nx;
                const float zs = ((k - g.nz / 2.0f) * g.ds) / g.nx;
                g.at(4, i, j, k) += 0.1f * std::exp(-(xs * xs + ys * ys + zs * zs) / (2.0f * sigma * sigma));
            }
}

// init_relaxation_3d (init.rs:726-732)
void init_relaxation(Grid &g)
{
    g.fill(0, kDensity);
    g.fill(4, kPressure);
}

// init_shock_tube (init.rs:734-757)
void init_shock_tube(Grid &g)
{
    const float length = g.nx * g.ds;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const bool left = i * g.ds < length / 2.0f;
                g.at(4, i, j, k) = left ? 1.0f : 0.1f;
                g.at(0, i, j, k) = left ? 1.0f : 0.125f;
            }
}

// hydrostatic column of the rising-bubble demos (init.rs:759-790)
void column(Grid &g)
{
    const float p0 = 101000.0f, t0 = 293.0f;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float zs = (g.nz - 2.0f * 2) * g.ds + (0.0f * g.ds) - (k - 2.0f) * g.ds;
                g.at(4, i, j, k) = p0 + std::fabs(kG) * zs;
                g.at(0, i, j, k) = g.at(4, i, j, k) / (kRs * t0);
            }
    hydrostatic_floor(g);
}

void bubble(Grid &g, const float c[3], float r_out, float r_in)
{
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float mag = std::sqrt(std::pow(i - c[0], 2.0f) + std::pow(j - c[1], 2.0f) + std::pow(k - c[2], 2.0f));
                if (mag < r_out) g.at(0, i, j, k) = 1.0f;
                if (mag < r_in) g.at(0, i, j, k) = 0.8f;
            }
}

// init_trb_demo (init.rs:759-812)
void init_trb(Grid &g)
{
    column(g);
    const float c[3] = {75.0f, 1.0f, 128.0f};
    bubble(g, c, 25.0f, 20.0f);
}

// init_trbs_demo (init.rs:814-896)
void init_trbs(Grid &g)
{
    column(g);
    const float c1[3] = {32.0f + 8.0f, 1.0f, 32.0f}, c2[3] = {88.0f - 8.0f, 1.0f, 32.0f}, c3[3] = {64.0f, 1.0f, 100.0f};
    bubble(g, c1, 15.0f, 10.0f);
    bubble(g, c2, 15.0f, 10.0f);
    bubble(g, c3, 1.5f * 15.0f, 1.5f * 10.0f);
}

// Benchmark configurations (gamma 1.4, CFL 0.4, no gravity, unit cube,
// cell centres x_i = (i - 2 + 1/2) ds): the Sod tube, the blast wave and the
// Kelvin-Helmholtz layer the benchmark is quoted on.
float centre(int i, float ds) { return ((float)i - 2.0f + 0.5f) * ds; }

void init_sod(Grid &g)
{
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const bool left = centre(i, g.ds) < 0.5f;
                g.at(0, i, j, k) = left ? 1.0f : 0.125f;
                g.at(4, i, j, k) = left ? 1.0f : 0.1f;
            }
}

void init_blast(Grid &g)
{
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const float x = centre(i, g.ds) - 0.5f, y = centre(j, g.ds) - 0.5f, z = centre(k, g.ds) - 0.5f;
                g.at(0, i, j, k) = 1.0f;
                g.at(4, i, j, k) = x * x + y * y + z * z < 0.1f * 0.1f ? 10.0f : 0.1f;
            }
}

// v perturbation: 0.01 sin(4 pi x) plus uniform noise in [-1e-3, 1e-3] from a
// seeded splitmix64 stream (the same distribution as the Python set-up, not
// the same numbers)
std::function<void(Grid &)> init_khi_layer(uint64_t seed)
{
    return [seed](Grid &g) {
        uint64_t st = seed;
        auto uni = [&st]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int i = 0; i < g.nx; ++i) {
            const float sx = 0.01f * std::sin(4.0f * kPi * centre(i, g.ds));
            for (int j = 0; j < g.ny; ++j) {
                const bool inner = std::fabs(centre(j, g.ds) - 0.5f) < 0.25f;
                for (int k = 0; k < g.nz; ++k) {
                    g.at(0, i, j, k) = inner ? 2.0f : 1.0f;
                    g.at(1, i, j, k) = inner ? 0.5f : -0.5f;
                    g.at(2, i, j, k) = sx + (float)(2e-3 * uni() - 1e-3);
                    g.at(4, i, j, k) = 2.5f;
                }
            }
        }
    };
}

bool find_config(const std::string &name, uint64_t seed, Demo &d, std::string &precision)
{
    if (name == "sod64") {
        d = {{64, 64, 64}, 1.0f / 64, {O, O, O, O, O, O}, false, init_sod};
        precision = "fp64";
    } else if (name == "blast256") {
        d = {{256, 256, 256}, 1.0f / 256, {P, P, P, P, P, P}, false, init_blast};
        precision = "fp64";
    } else if (name == "khi512") {
        d = {{512, 512, 512}, 1.0f / 512, {P, P, P, P, P, P}, false, init_khi_layer(seed)};
        precision = "fp32";
    } else {
        return false;
    }
    return true;
}

// set_up_*_demo (init.rs:8-321) and the dispatch of main.rs:79-108
bool find_demo(const std::string &name, Demo &d)
{
    if (name == "KELVIN_HELMHOLTZ") d = {{256, 2, 128}, 1.0f / 256, {P, P, P, P, P, P}, false, init_khi};
    else if (name == "RAYLEIGH_TAYLOR") d = {{100, 2, 300}, 2.0f * 0.0025f, {P, P, P, P, W, W}, true, init_rt};
    else if (name == "RAYLEIGH_BERNARD") d = {{64, 64, 64}, 1.0f / 64, {W, W, W, W, W, W}, true, init_rb};
    else if (name == "EXPLOSION") d = {{128, 2, 128}, 1.0f / 128, {W, W, P, P, W, W}, false, init_explosion_2d};
    else if (name == "RELAXATION") d = {{30, 30, 30}, 0.1f, {W, W, W, W, W, O}, true, init_relaxation};
    else if (name == "SHOCK_TUBE") d = {{5000, 2, 2}, 0.0002f, {W, W, P, P, P, P}, false, init_shock_tube};
    else if (name == "THERMAL_RISING_BUBBLE") d = {{150, 2, 512}, 1.0f, {W, W, P, P, W, O}, true, init_trb};
    else if (name == "THERMAL_RISING_BUBBLES") d = {{128, 2, 500}, 2.0f, {W, W, P, P, W, O}, true, init_trbs};
    else return false;
    return true;
}

bool write_npy(const std::string &path, const std::vector<double> &v, const std::vector<int> &shape)
{
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    std::string dims;
    for (size_t i = 0; i < shape.size(); ++i) dims += std::to_string(shape[i]) + ", ";
    const std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + dims + "), }";
    const int total = 10 + (int)hdr.size() + 1;
    const int pad = (64 - total % 64) % 64;
    std::string h = hdr;
    h.append(pad, ' ');
    h.push_back('\n');
    const unsigned short hl = (unsigned short)h.size();
    std::fwrite("\x93NUMPY\x01\x00", 1, 8, fp);
    std::fwrite(&hl, 2, 1, fp);
    std::fwrite(h.data(), 1, h.size(), fp);
    const bool ok = std::fwrite(v.data(), sizeof(double), v.size(), fp) == v.size();
    return std::fclose(fp) == 0 && ok;
}

// A float64 C-order .npy file written by write_npy; false unless its shape is `shape`.
bool read_npy(const std::string &path, const std::vector<int> &shape, std::vector<double> &v)
{
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    char magic[8];
    unsigned short hl = 0;
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, "\x93NUMPY\x01\x00", 8) == 0 &&
              std::fread(&hl, 2, 1, fp) == 1;
    std::string h(hl, ' ');
    ok = ok && std::fread(&h[0], 1, hl, fp) == hl;
    std::string dims;
    for (size_t i = 0; i < shape.size(); ++i) dims += std::to_string(shape[i]) + ", ";
    ok = ok && h.find("'descr': '<f8'") != std::string::npos && h.find("'fortran_order': False") != std::string::npos &&
         h.find("'shape': (" + dims + ")") != std::string::npos;
    size_t n = 1;
    for (int d : shape) n *= (size_t)d;
    v.assign(n, 0.0);
    ok = ok && std::fread(v.data(), sizeof(double), n, fp) == n;
    std::fclose(fp);
    return ok;
}

struct Clock {
    long long step = 0, wstep = 0;
    float tnow = 0.0f, wtime = 0.0f;
};

const char *kPrimNames[5] = {"rho", "vx", "vy", "vz", "pressure"};
const char *kConsNames[5] = {"mass", "momentum_x", "momentum_y", "momentum_z", "energy"};

// host arrays of the library's element type, as float64 for the .npy files
std::vector<double> to_f64(const std::vector<char> &b, size_t n, size_t el)
{
    std::vector<double> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = el == 8 ? reinterpret_cast<const double *>(b.data())[i] : (double)reinterpret_cast<const float *>(b.data())[i];
    return v;
}

// One step whose result is dumped: save_to_disk (io.rs:5-196) after
// compute_step (main.rs:56-62) -- the four groups, f64, ghosted -- plus the
// loop clock after the step (`after`) for --restart.
int step_and_dump(dg_ctx *ctx, const Grid &shape, int precision, const std::string &dir, double *dt,
                  const Clock &before)
{
    mkdir(dir.c_str(), 0755);
    const size_t n = (size_t)shape.nx * shape.ny * shape.nz;
    const size_t el = precision == DG_FP64 ? 8 : 4;
    std::vector<std::vector<char>> prim(5, std::vector<char>(n * el)), grad(5, std::vector<char>(3 * n * el)),
        flux(5, std::vector<char>(3 * n * el)), cons(5, std::vector<char>(n * el));
    void *pp[5], *pg[5], *pf[5], *pc[5];
    for (int q = 0; q < 5; ++q) {
        pp[q] = prim[q].data();
        pg[q] = grad[q].data();
        pf[q] = flux[q].data();
        pc[q] = cons[q].data();
    }
    int rc = dg_step_dump(ctx, dt, pp, pg, pf);
    if (!rc) rc = dg_get_state(ctx, pc, DG_CONSERVED);
    if (rc) return rc;
    const std::vector<int> s3 = {shape.nx, shape.ny, shape.nz}, s4 = {shape.nx, shape.ny, shape.nz, 3};
    for (int q = 0; q < 5; ++q) {
        if (!write_npy(dir + "/primitives." + kPrimNames[q] + ".npy", to_f64(prim[q], n, el), s3) ||
            !write_npy(dir + "/conserved." + kConsNames[q] + ".npy", to_f64(cons[q], n, el), s3) ||
            !write_npy(dir + "/gradients." + kPrimNames[q] + ".npy", to_f64(grad[q], 3 * n, el), s4) ||
            !write_npy(dir + "/fluxes." + kConsNames[q] + ".npy", to_f64(flux[q], 3 * n, el), s4))
            return DG_ERR_INVALID;
    }
    // the clock at the top of the next loop iteration: the dump resets wtime
    // and counts wstep, then tnow += dt; wtime += dt; step += 1 (main.rs:61-66)
    Clock after = before;
    after.wtime = 0.0f;
    after.wstep += 1;
    after.tnow += (float)*dt;
    after.wtime += (float)*dt;
    after.step += 1;
    FILE *fp = std::fopen((dir + "/meta.txt").c_str(), "w");
    if (!fp) return DG_ERR_INVALID;
    std::fprintf(fp, "step %lld\ntnow %.9g\nwtime %.9g\nwstep %lld\n", after.step, (double)after.tnow,
                 (double)after.wtime, after.wstep);
    return std::fclose(fp) == 0 ? DG_OK : DG_ERR_INVALID;
}

// --restart: the conserved state and the loop clock of a dump.
int load_dump(dg_ctx *ctx, const Grid &shape, int precision, const std::string &dir, Clock &clk)
{
    const size_t n = (size_t)shape.nx * shape.ny * shape.nz;
    const size_t el = precision == DG_FP64 ? 8 : 4;
    const std::vector<int> s3 = {shape.nx, shape.ny, shape.nz};
    std::vector<std::vector<char>> buf(5, std::vector<char>(n * el));
    const void *f[5];
    for (int q = 0; q < 5; ++q) {
        std::vector<double> v;
        if (!read_npy(dir + "/conserved." + kConsNames[q] + ".npy", s3, v)) {
            std::fprintf(stderr, "cannot read %s/conserved.%s.npy with shape (%d, %d, %d)\n", dir.c_str(),
                         kConsNames[q], shape.nx, shape.ny, shape.nz);
            return DG_ERR_INVALID;
        }
        for (size_t i = 0; i < n; ++i) {
            if (el == 8) reinterpret_cast<double *>(buf[q].data())[i] = v[i];
            else reinterpret_cast<float *>(buf[q].data())[i] = (float)v[i];
        }
        f[q] = buf[q].data();
    }
    FILE *fp = std::fopen((dir + "/meta.txt").c_str(), "r");
    double tnow = 0, wtime = 0;
    const bool ok = fp && std::fscanf(fp, "step %lld\ntnow %lf\nwtime %lf\nwstep %lld", &clk.step, &tnow, &wtime,
                                      &clk.wstep) == 4;
    if (fp) std::fclose(fp);
    if (!ok) {
        std::fprintf(stderr, "cannot read %s/meta.txt\n", dir.c_str());
        return DG_ERR_INVALID;
    }
    clk.tnow = (float)tnow;
    clk.wtime = (float)wtime;
    return dg_set_state(ctx, f, DG_CONSERVED);
}

// ------------------------------------------------ ensemble and learned --
// Seeded splitmix64 streams for the two workloads below: the same
// distributions as the Python set-ups (deepgreen/setups.py), not the same
// numbers.
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uni(); }
    double normal()
    {
        const double u1 = (double)((next() >> 11) + 1) * (1.0 / 9007199254740993.0), u2 = uni();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Host primitive fields of a ghosted grid in the context's precision.
struct HostState {
    std::vector<std::vector<char>> buf;
    const void *f[5];
    HostState(const Grid &g, bool fp64) : buf(5)
    {
        const size_t n = (size_t)g.nx * g.ny * g.nz;
        for (int q = 0; q < 5; ++q) {
            buf[q].resize(n * (fp64 ? 8 : 4));
            for (size_t i = 0; i < n; ++i) {
                if (fp64) reinterpret_cast<double *>(buf[q].data())[i] = g.f[q][i];
                else reinterpret_cast<float *>(buf[q].data())[i] = g.f[q][i];
            }
            f[q] = buf[q].data();
        }
    }
};

// Cell centre of ghosted index i on the unit cube with n interior cells.
double unit_centre(int i, int n) { return (i - 2 + 0.5) / n; }

// Random planar Riemann problem (the ensemble's members): a normal uniform on
// the sphere through the cube's centre; on each side rho, p ~ U[0.1, 1] and
// each velocity component ~ U[-1, 1].
void init_planar_riemann(Grid &g, int n, Rng &r)
{
    double nrm[3] = {r.normal(), r.normal(), r.normal()};
    const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (double &v : nrm) v /= len;
    float side[2][5];
    for (auto &sd : side) {
        sd[0] = (float)r.uniform(0.1, 1.0);
        sd[4] = (float)r.uniform(0.1, 1.0);
        for (int d = 1; d <= 3; ++d) sd[d] = (float)r.uniform(-1.0, 1.0);
    }
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const double s = (unit_centre(i, n) - 0.5) * nrm[0] + (unit_centre(j, n) - 0.5) * nrm[1] +
                                 (unit_centre(k, n) - 0.5) * nrm[2];
                const float *v = side[s < 0.0 ? 0 : 1];
                for (int q = 0; q < 5; ++q) g.at(q, i, j, k) = v[q];
            }
}

// The learned workload's initial condition on an n^3 periodic unit cube:
// density 1 + 8 Fourier modes (integer wave vectors in [-4, 4]^3, amplitudes
// in [-0.2, 0.2], random phases), scaled down uniformly where they would take
// the density below 0.2; pressure 1 + 1-3 Gaussian blasts of peak 5-20 (width
// 0.05, periodic distance); at rest.
void init_learned(Grid &g, int n, uint64_t seed)
{
    Rng r{seed};
    struct Mode { double k[3], a, ph; };
    std::vector<Mode> modes;
    while (modes.size() < 8) {
        Mode m;
        bool any = false;
        for (double &k : m.k) {
            k = (double)(int)(r.next() % 9) - 4.0;
            any = any || k != 0.0;
        }
        if (!any) continue;
        m.a = r.uniform(-0.2, 0.2);
        m.ph = r.uniform(0.0, 2.0 * M_PI);
        modes.push_back(m);
    }
    struct Blast { double c[3], peak; };
    std::vector<Blast> blasts(1 + r.next() % 3);
    for (Blast &b : blasts) {
        for (double &c : b.c) c = r.uni();
        b.peak = r.uniform(5.0, 20.0);
    }
    const double sigma = 0.05;
    auto gauss = [&](double x, double c) {
        double d = std::fabs(x - c);
        d = std::min(d, 1.0 - d);
        return std::exp(-d * d / (2.0 * sigma * sigma));
    };
    std::vector<double> S((size_t)g.nx * g.ny * g.nz);
    double smin = 0.0;
    for (int i = 0; i < g.nx; ++i)
        for (int j = 0; j < g.ny; ++j)
            for (int k = 0; k < g.nz; ++k) {
                const double x = unit_centre(i, n), y = unit_centre(j, n), z = unit_centre(k, n);
                double sm = 0.0, p = 1.0;
                for (const Mode &m : modes) sm += m.a * std::sin(2.0 * M_PI * (m.k[0] * x + m.k[1] * y + m.k[2] * z) + m.ph);
                for (const Blast &b : blasts) p += (b.peak - 1.0) * gauss(x, b.c[0]) * gauss(y, b.c[1]) * gauss(z, b.c[2]);
                const size_t c = ((size_t)i * g.ny + j) * g.nz + k;
                S[c] = sm;
                smin = std::min(smin, sm);
                g.f[4][c] = (float)p;
            }
    const double rho_min = 0.2, scale = 1.0 + smin >= rho_min ? 1.0 : (1.0 - rho_min) / (-smin);
    for (size_t c = 0; c < S.size(); ++c) g.f[0][c] = (float)(1.0 + scale * S[c]);
}

// The learned workload's solver: periodic, FP32 fast math, minmod, gamma 1.4, CFL 0.4.
dg_config learned_config(int n)
{
    dg_config cfg;
    dg_config_default(&cfg);
    cfg.gamma = 1.4;
    cfg.cfl = 0.4;
    cfg.g = 0.0;
    for (int d = 0; d < 3; ++d) cfg.n[d] = n;
    for (int d = 0; d < 6; ++d) cfg.bc[d] = DG_PERIODIC;
    cfg.ds = 1.0 / n;
    cfg.precision = DG_FP32;
    cfg.math = DG_MATH_FAST;
    cfg.limiter = DG_LIMITER_MINMOD;
    return cfg;
}

// --config ensemble: `members` cubic grids with edges drawn by seed from {32,
// 48, 64, 96, 128}, random planar Riemann problems, outflow, gamma 1.4, CFL
// 0.4, advanced `steps` steps in one batched run.
int run_ensemble(uint64_t seed, int members, long long steps, int prec, int math, int limiter)
{
    static const int sizes[5] = {32, 48, 64, 96, 128};
    Rng r{seed};
    std::vector<dg_config> cfgs(members);
    long long cells = 0;
    for (dg_config &c : cfgs) {
        dg_config_default(&c);
        const int n = sizes[r.next() % 5];
        c.gamma = 1.4;
        c.cfl = 0.4;
        c.g = 0.0;
        for (int d = 0; d < 3; ++d) c.n[d] = n;
        for (int d = 0; d < 6; ++d) c.bc[d] = DG_OUTFLOW;
        c.ds = 1.0 / n;
        c.precision = prec;
        c.math = math;
        c.limiter = limiter;
        cells += (long long)n * n * n;
    }
    dg_ensemble *e = nullptr;
    int rc = dg_ensemble_create(cfgs.data(), members, 0, &e);
    if (rc) {
        std::fprintf(stderr, "dg_ensemble_create failed (%d): %s\n", rc, dg_ensemble_last_error(e));
        return 1;
    }
    std::printf("Running ensemble of %d grids (%lld cells).\n", members, cells);
    for (int i = 0; i < members; ++i) {
        const int n = cfgs[i].n[0];
        Grid g(n, n, n, (float)cfgs[i].ds);
        Rng ri{seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i + 1};
        init_planar_riemann(g, n, ri);
        HostState h(g, prec == DG_FP64);
        if ((rc = dg_ensemble_set_state(e, i, h.f, DG_PRIMITIVE))) {
            std::fprintf(stderr, "dg_ensemble_set_state: %s\n", dg_ensemble_last_error(e));
            return 1;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    rc = dg_ensemble_run(e, steps, INFINITY);
    double t = 0, nd = 0;
    int64_t done = 0;
    if (!rc) rc = dg_ensemble_clock(e, 0, &t, &done, &nd);   // waits for the run
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
        std::fprintf(stderr, "dg_ensemble_run: %s\n", dg_ensemble_last_error(e));
        return 1;
    }
    bool finite = true;
    for (int i = 0; i < members; ++i) {
        double sums[5];
        finite = finite && dg_ensemble_conservation_sums(e, i, sums) == 0;
        for (double v : sums) finite = finite && std::isfinite(v);
    }
    dg_ensemble_destroy(e);
    std::printf("Ensemble done in %.3fs! (%lld steps, %.2f Gcell-updates/s including graph capture, state %s)\n",
                secs, (long long)done, cells * (double)done / secs / 1e9, finite ? "finite" : "NOT FINITE");
    return finite ? 0 : 1;
}

// --config learned128: the learned flux correction on a 128^3 grid of the
// workload's initial condition (seed 1000 + seed).  The network's parameters
// come from --params (float32: dg_cnn_param_count() parameters, then the 15
// normalization values, as FluxCNN.get_params / get_normalization give
// them), or are trained here on `train_ics` 256^3 plain runs (a sample every
// 4 fine steps up to t = 0.03, then `train_steps` Adam steps, lr 1e-3, batch
// 8 of 32^3 crops).
int run_learned(uint64_t seed, long long steps, const std::string &params, int train_ics, int train_steps)
{
    dg_cnn *m = nullptr;
    int rc = dg_cnn_create(0, seed, &m);
    if (rc) { std::fprintf(stderr, "dg_cnn_create failed (%d)\n", rc); return 1; }
    const int64_t np = dg_cnn_param_count();
    if (!params.empty()) {
        std::vector<float> p(np + 15);
        FILE *fp = std::fopen(params.c_str(), "rb");
        const bool ok = fp && std::fread(p.data(), sizeof(float), p.size(), fp) == p.size();
        if (fp) std::fclose(fp);
        if (!ok || dg_cnn_set_params(m, p.data()) || dg_cnn_set_normalization(m, p.data() + np)) {
            std::fprintf(stderr, "cannot read %lld parameters and 15 normalization values from %s\n", (long long)np,
                         params.c_str());
            return 1;
        }
        std::printf("Network parameters from %s.\n", params.c_str());
    } else if (train_ics > 0) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 1; k <= train_ics; ++k) {
            dg_config cf = learned_config(256), cc = learned_config(128);
            dg_ctx *fine = nullptr, *coarse = nullptr;
            if ((rc = dg_create(&cf, &fine)) || (rc = dg_create(&cc, &coarse))) {
                std::fprintf(stderr, "dg_create failed (%d)\n", rc);
                return 1;
            }
            Grid g(256, 256, 256, (float)cf.ds);
            init_learned(g, 256, seed + (uint64_t)k);
            HostState h(g, false);
            if ((rc = dg_set_state(fine, h.f, DG_PRIMITIVE))) return 1;
            double t = 0;
            int64_t n = 0;
            while (t < 0.03) {
                if ((rc = dg_cnn_add_sample(m, fine, coarse)) || (rc = dg_run(fine, 2, INFINITY, &t, &n))) {
                    std::fprintf(stderr, "sampling failed (%d)\n", rc);
                    return 1;
                }
            }
            dg_destroy(fine);
            dg_destroy(coarse);
            std::printf("Training run %d of %d: %d samples.\n", k, train_ics, dg_cnn_sample_count(m));
            std::fflush(stdout);
        }
        std::vector<float> losses(train_steps);
        if ((rc = dg_cnn_fit_normalization(m)) ||
            (rc = dg_cnn_train(m, train_steps, 8, 32, 1e-3f, seed, losses.data()))) {
            std::fprintf(stderr, "training failed (%d)\n", rc);
            return 1;
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("Trained on %d samples in %.1fs, loss %g -> %g.\n", dg_cnn_sample_count(m), secs,
                    (double)losses.front(), (double)losses.back());
    } else {
        std::printf("Untrained network (seeded weights).\n");
    }
    dg_config cfg = learned_config(128);
    dg_ctx *ctx = nullptr;
    if ((rc = dg_create(&cfg, &ctx))) { std::fprintf(stderr, "dg_create failed (%d)\n", rc); return 1; }
    Grid g(128, 128, 128, (float)cfg.ds);
    init_learned(g, 128, 1000 + seed);
    HostState h(g, false);
    if ((rc = dg_set_state(ctx, h.f, DG_PRIMITIVE))) return 1;
    double s0[5], s1[5];
    dg_conservation_sums(ctx, s0);
    double t = 0;
    int64_t n = 0;
    const auto t0 = std::chrono::steady_clock::now();
    rc = dg_learned_run(ctx, m, steps, INFINITY, &t, &n);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) { std::fprintf(stderr, "dg_learned_run: %s\n", dg_last_error(ctx)); return 1; }
    int64_t flagged = 0;
    dg_learned_flagged(m, &flagged);
    dg_conservation_sums(ctx, s1);
    const bool finite = std::isfinite(s1[0]) && std::isfinite(s1[4]);
    std::printf("Learned run done in %.3fs! (%lld steps, t = %g, mass drift %.2e, %lld cells on the plain flux in "
                "the last step, state %s)\n", secs, (long long)n, t, (s1[0] - s0[0]) / s0[0], (long long)flagged,
                finite ? "finite" : "NOT FINITE");
    dg_destroy(ctx);
    dg_cnn_destroy(m);
    return finite ? 0 : 1;
}

void usage()
{
    std::fprintf(stderr,
                 "usage: dg_cli --demo NAME --tout T --total-time T --max-steps N\n"
                 "              [--precision fp32|fp64|fp32_bulk] [--math exact|fast]\n"
                 "              [--limiter none|minmod|vanleer] [--out DIR] [--no-dump] [--quiet]\n"
                 "              [--restart DIR/state.NNNNN] (continue from a dump of the same set-up)\n"
                 "       dg_cli --config sod64|blast256|khi512 --steps N [--seed S] [options above]\n"
                 "              (gamma 1.4, CFL 0.4, no gravity; dumps only with --out: after the\n"
                 "              first step, then every --tout; default precision fp64, fp64, fp32)\n"
                 "       dg_cli --config ensemble --steps N [--seed S] [--members M]\n"
                 "              [--precision ..] [--math ..] [--limiter ..] (default fp32, fast, minmod)\n"
                 "       dg_cli --config learned128 --steps N [--seed S] [--params FILE]\n"
                 "              [--train-ics K] [--train-steps T] (default 32 and 2000)\n"
                 "demos: KELVIN_HELMHOLTZ THERMAL_RISING_BUBBLE THERMAL_RISING_BUBBLES EXPLOSION RELAXATION\n"
                 "       SHOCK_TUBE RAYLEIGH_TAYLOR RAYLEIGH_BERNARD\n");
}

}  // namespace

int main(int argc, char **argv)
{
    std::string demo, config, out = ".", precision, math = "exact", limiter = "none", restart;
    float tout = -1, total_time = -1;
    long long max_steps = -1;
    uint64_t seed = 1;
    bool do_dump = true, quiet = false, out_given = false, math_given = false, limiter_given = false;
    int members = 256, train_ics = 32, train_steps = 2000;
    std::string params;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        auto val = [&]() -> std::string {
            if (a + 1 >= argc) { usage(); std::exit(2); }
            return argv[++a];
        };
        if (k == "--demo" || k == "-d") demo = val();
        else if (k == "--tout" || k == "-t") tout = std::strtof(val().c_str(), nullptr);
        else if (k == "--total-time") total_time = std::strtof(val().c_str(), nullptr);
        else if (k == "--max-steps" || k == "-m") max_steps = std::strtoll(val().c_str(), nullptr, 10);
        else if (k == "--precision") precision = val();
        else if (k == "--math") { math = val(); math_given = true; }
        else if (k == "--limiter") { limiter = val(); limiter_given = true; }
        else if (k == "--members") members = std::atoi(val().c_str());
        else if (k == "--params") params = val();
        else if (k == "--train-ics") train_ics = std::atoi(val().c_str());
        else if (k == "--train-steps") train_steps = std::atoi(val().c_str());
        else if (k == "--out") { out = val(); out_given = true; }
        else if (k == "--no-dump") do_dump = false;
        else if (k == "--quiet") quiet = true;
        else if (k == "--config") config = val();
        else if (k == "--steps") max_steps = std::strtoll(val().c_str(), nullptr, 10);
        else if (k == "--seed") seed = std::strtoull(val().c_str(), nullptr, 10);
        else if (k == "--restart") restart = val();
        else { usage(); return 2; }
    }
    const bool known = (precision.empty() || precision == "fp32" || precision == "fp64" || precision == "fp32_bulk") &&
                       (math == "exact" || math == "fast") &&
                       (limiter == "none" || limiter == "minmod" || limiter == "vanleer");
    if (!known) {
        std::fprintf(stderr, "unknown --precision, --math or --limiter value\n");
        usage();
        return 2;
    }
    if (config == "ensemble" || config == "learned128") {
        if (!demo.empty() || max_steps < 0 || members < 1 || train_ics < 0 || train_steps < 1) { usage(); return 2; }
        if (config == "learned128") return run_learned(seed, max_steps, params, train_ics, train_steps);
        const int prec = precision == "fp64" ? DG_FP64 : (precision == "fp32_bulk" ? DG_FP32_BULK : DG_FP32);
        const int mth = !math_given || math == "fast" ? DG_MATH_FAST : DG_MATH_EXACT;
        const int lim = !limiter_given || limiter == "minmod" ? DG_LIMITER_MINMOD
                                                              : (limiter == "vanleer" ? DG_LIMITER_VANLEER : DG_LIMITER_NONE);
        return run_ensemble(seed, members, max_steps, prec, mth, lim);
    }
    Demo d;
    std::string default_precision = "fp32";
    if (!config.empty()) {
        if (!demo.empty() || max_steps < 0) { usage(); return 2; }
        if (!find_config(config, seed, d, default_precision)) {
            std::fprintf(stderr, "Configuration not found!\n");
            usage();
            return 2;
        }
        demo = config;
        if (total_time < 0) total_time = INFINITY;
        if (tout < 0) tout = INFINITY;
        if (!out_given) do_dump = false;   // a 512^3 dump is 11 GB: only when asked for
    } else {
        if (demo.empty() || tout < 0 || total_time < 0 || max_steps < 0) { usage(); return 2; }
        if (!find_demo(demo, d)) {
            std::fprintf(stderr, "Demo not found! Available demos are:\n");
            usage();
            return 0;   // the reference exits 0 here too (main.rs:103-108)
        }
    }
    if (precision.empty()) precision = default_precision;
    const auto t0 = std::chrono::steady_clock::now();
    dg_config cfg;
    dg_config_default(&cfg);   // GAMMA, CFL, G of globals.rs
    if (!config.empty()) {
        cfg.gamma = 1.4;
        cfg.cfl = 0.4;
        cfg.g = 0.0;
    }
    for (int i = 0; i < 3; ++i) cfg.n[i] = d.n[i];
    for (int i = 0; i < 6; ++i) cfg.bc[i] = d.bc[i];
    cfg.ds = d.ds;
    cfg.gravity = d.gravity;
    cfg.precision = precision == "fp64" ? DG_FP64 : (precision == "fp32_bulk" ? DG_FP32_BULK : DG_FP32);
    cfg.math = math == "fast" ? DG_MATH_FAST : DG_MATH_EXACT;
    cfg.limiter = limiter == "minmod" ? DG_LIMITER_MINMOD : (limiter == "vanleer" ? DG_LIMITER_VANLEER : DG_LIMITER_NONE);
    dg_ctx *ctx = nullptr;
    int rc = dg_create(&cfg, &ctx);
    if (rc) { std::fprintf(stderr, "dg_create failed (%d)\n", rc); return 1; }
    Grid g(d.n[0], d.n[1], d.n[2], d.ds);
    Clock clk;
    if (!restart.empty()) {
        // continue a run from one of its dumps (state and loop clock)
        if ((rc = load_dump(ctx, g, cfg.precision, restart, clk))) {
            std::fprintf(stderr, "restart from %s failed (%d)\n", restart.c_str(), rc);
            return 1;
        }
        std::printf("Restarting simulation %s from %s: step %lld, time %g.\n", demo.c_str(), restart.c_str(),
                    clk.step, (double)clk.tnow);
    } else {
        d.init(g);
        std::printf("Running simulation %s: %dx%dx%d cells, ds %g.\n", demo.c_str(), d.n[0], d.n[1], d.n[2],
                    (double)d.ds);
        // state upload (calc_conserved of main.rs:50 happens in the library)
        std::vector<std::vector<char>> host(5);
        const size_t n = (size_t)g.nx * g.ny * g.nz, el = cfg.precision == DG_FP64 ? 8 : 4;
        const void *f[5];
        for (int q = 0; q < 5; ++q) {
            host[q].resize(n * el);
            for (size_t i = 0; i < n; ++i) {
                if (el == 8) reinterpret_cast<double *>(host[q].data())[i] = g.f[q][i];
                else reinterpret_cast<float *>(host[q].data())[i] = g.f[q][i];
            }
            f[q] = host[q].data();
        }
        if ((rc = dg_set_state(ctx, f, DG_PRIMITIVE))) {
            std::fprintf(stderr, "dg_set_state: %s\n", dg_last_error(ctx));
            return 1;
        }
    }
    // compute() of main.rs:38-68 (dt starts at 0.5 only for the first print)
    float dt = 0.5f;
    while (!(clk.tnow > total_time || clk.step >= max_steps)) {
        if (!quiet) std::printf("Computing step %lld and time %g with dt %g. \n", clk.step, (double)clk.tnow, (double)dt);
        double dtd = 0;
        // the dump test of main.rs:57 depends only on wtime and wstep, which
        // the step does not change: a dumped step is taken by dg_step_dump
        const bool dump_now = clk.wtime > tout || clk.wstep == 0;
        if (dump_now && do_dump) {
            char name[32];
            std::snprintf(name, sizeof(name), "state.%05lld", clk.wstep);
            if ((rc = step_and_dump(ctx, g, cfg.precision, out + "/" + name, &dtd, clk))) {
                std::fprintf(stderr, "dump failed (%d): %s\n", rc, dg_last_error(ctx));
                return 1;
            }
        } else if ((rc = dg_step(ctx, &dtd))) {
            std::fprintf(stderr, "dg_step: %s\n", dg_last_error(ctx));
            return 1;
        }
        dt = (float)dtd;
        if (dump_now) {
            clk.wtime = 0.0f;
            clk.wstep += 1;
        }
        clk.tnow += dt;
        clk.wtime += dt;
        clk.step += 1;
    }
    dg_destroy(ctx);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation done in %.3fs! (%lld steps, t = %g, %lld dumps)\n", secs, clk.step, (double)clk.tnow,
                clk.wstep);
    return 0;
}
