// oracle_capi.cpp -- C entry points of the CPU ORACLE (test infrastructure
// only; loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg through ctypes).  Wraps mpm_oracle::Solver<float> (FP32
// restatement of the reference) and Solver<double> (FP64 rerun).
#include <chrono>
#include <cstdio>
#include <memory>
#include <new>

#include "mpm_oracle.hpp"

using namespace mpm_oracle;

namespace {

struct Handle {
    int precision;  // 32 or 64
    std::unique_ptr<Solver<float>> s32;
    std::unique_ptr<Solver<double>> s64;
};

template <class R>
std::vector<Particle<R>> from_input(const mpm_particle_input* in) {
    std::vector<Particle<R>> P;
    P.reserve(size_t(in->count));
    for (int64_t i = 0; i < in->count; ++i)
        P.push_back(make_particle<R>(in->position + 3 * i, in->velocity + 3 * i, in->mass[i], in->hardening[i],
                                     in->young[i], in->poisson[i], in->critical_compression[i],
                                     in->critical_stretch[i], in->model[i]));
    return P;
}

template <class R>
std::vector<Particle<R>> from_state(const mpm_particle_state* st) {
    std::vector<Particle<R>> P(size_t(st->count));
    for (int64_t i = 0; i < st->count; ++i) {
        Particle<R>& p = P[size_t(i)];
        p.volume0 = st->volume0[i];
        p.mass = st->mass[i];
        for (int a = 0; a < 3; ++a) {
            p.position[a] = st->position[3 * i + a];
            p.velocity[a] = st->velocity[3 * i + a];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                p.Fe[r][c] = st->Fe[9 * i + 3 * r + c];
                p.Fp[r][c] = st->Fp[9 * i + 3 * r + c];
                p.B[r][c] = st->B ? st->B[9 * i + 3 * r + c] : R(0);
            }
        p.mu0 = st->mu0[i];
        p.lambda0 = st->lambda0[i];
        p.hardening = st->hardening[i];
        p.crit_c = st->critical_compression[i];
        p.crit_s = st->critical_stretch[i];
        p.model = st->model[i];
        p.D = mzero<R>();
        p.Dinv = mzero<R>();
    }
    return P;
}

template <class R>
void set_apic_d(Solver<R>& s) {
    // D / Dinv are set by initialize() in the reference; a restored state
    // without initialize still needs them (src/apic_mpm_solver3d.cu:151-157).
    if (!s.apic) return;
    R k = 0;
    if (s.desc.interp == MPM_INTERP_CUBIC_BSPLINE) k = R(1 / 3.0);
    if (s.desc.interp == MPM_INTERP_QUADRATIC_BSPLINE) k = R(1 / 4.0);
    for (auto& p : s.P) {
        p.D = mscale(k * s.h * s.h, eye<R>());
        p.Dinv = inv(p.D);
    }
}

template <class R>
void to_state(const Solver<R>& s, mpm_particle_state* out) {
    for (int64_t i = 0; i < out->count && i < int64_t(s.P.size()); ++i) {
        const Particle<R>& p = s.P[size_t(i)];
        if (out->volume0) out->volume0[i] = float(p.volume0);
        if (out->mass) out->mass[i] = float(p.mass);
        for (int a = 0; a < 3; ++a) {
            if (out->position) out->position[3 * i + a] = float(p.position[a]);
            if (out->velocity) out->velocity[3 * i + a] = float(p.velocity[a]);
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                if (out->Fe) out->Fe[9 * i + 3 * r + c] = float(p.Fe[r][c]);
                if (out->Fp) out->Fp[9 * i + 3 * r + c] = float(p.Fp[r][c]);
                if (out->B) out->B[9 * i + 3 * r + c] = float(p.B[r][c]);
            }
        if (out->mu0) out->mu0[i] = float(p.mu0);
        if (out->lambda0) out->lambda0[i] = float(p.lambda0);
        if (out->hardening) out->hardening[i] = float(p.hardening);
        if (out->critical_compression) out->critical_compression[i] = float(p.crit_c);
        if (out->critical_stretch) out->critical_stretch[i] = float(p.crit_s);
        if (out->model) out->model[i] = p.model;
    }
}

// FP64 particle state read (for error reports against the FP64 rerun).
template <class R>
void to_state64(const Solver<R>& s, double* pos, double* vel, double* Fe, double* Fp, double* B) {
    for (size_t i = 0; i < s.P.size(); ++i) {
        const Particle<R>& p = s.P[i];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * i + a] = double(p.position[a]);
            if (vel) vel[3 * i + a] = double(p.velocity[a]);
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                if (Fe) Fe[9 * i + 3 * r + c] = double(p.Fe[r][c]);
                if (Fp) Fp[9 * i + 3 * r + c] = double(p.Fp[r][c]);
                if (B) B[9 * i + 3 * r + c] = double(p.B[r][c]);
            }
    }
}

template <class R>
void to_grid(const Solver<R>& s, mpm_grid_host* g) {
    for (size_t i = 0; i < s.G.size() && int64_t(i) < g->nodes; ++i) {
        const Node<R>& n = s.G[i];
        if (g->mass) g->mass[i] = float(n.mass);
        for (int a = 0; a < 3; ++a) {
            if (g->velocity) g->velocity[3 * i + a] = float(n.velocity[a]);
            if (g->velocity_star) g->velocity_star[3 * i + a] = float(n.velocity_star[a]);
            if (g->force) g->force[3 * i + a] = float(n.force[a]);
        }
    }
}

bool check_desc(const mpm_desc* d) {
    if (!d) return false;
    if (d->interp < 0 || d->interp > 2 || d->transfer < 0 || d->transfer > 1) return false;
    for (int i = 0; i < 3; ++i)
        if (d->resolution[i] <= 0) return false;
    return d->stride > 0.0f;
}

template <class F>
int dispatch(void* h, F&& f) {
    if (!h) return -1;
    Handle* H = static_cast<Handle*>(h);
    if (H->s32) f(*H->s32);
    else f(*H->s64);
    return 0;
}

}  // namespace

extern "C" {

// threads > 1 runs initialize() and every later step under OpenMP with atomic
// grid sums (the all-cores CPU baseline; sums in run-dependent order).
__attribute__((visibility("default"))) void* orc_create_threads(const mpm_desc* d, const mpm_particle_input* in,
                                                                int precision, int threads) {
    if (!check_desc(d) || !in || threads < 1) return nullptr;
    Handle* H = new (std::nothrow) Handle();
    if (!H) return nullptr;
    H->precision = precision;
    if (precision == 64) {
        H->s64.reset(new Solver<double>(*d, from_input<double>(in)));
        H->s64->threads = threads;
        H->s64->initialize();
    } else {
        H->s32.reset(new Solver<float>(*d, from_input<float>(in)));
        H->s32->threads = threads;
        H->s32->initialize();
    }
    return H;
}

__attribute__((visibility("default"))) void* orc_create(const mpm_desc* d, const mpm_particle_input* in,
                                                        int precision) {
    return orc_create_threads(d, in, precision, 1);
}

__attribute__((visibility("default"))) void* orc_create_state(const mpm_desc* d, const mpm_particle_state* st,
                                                              int run_initialize, int precision) {
    if (!check_desc(d) || !st) return nullptr;
    Handle* H = new (std::nothrow) Handle();
    if (!H) return nullptr;
    H->precision = precision;
    if (precision == 64) {
        H->s64.reset(new Solver<double>(*d, from_state<double>(st)));
        set_apic_d(*H->s64);
        if (run_initialize) H->s64->initialize();
    } else {
        H->s32.reset(new Solver<float>(*d, from_state<float>(st)));
        set_apic_d(*H->s32);
        if (run_initialize) H->s32->initialize();
    }
    return H;
}

__attribute__((visibility("default"))) void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

__attribute__((visibility("default"))) int orc_step(void* h, int n) {
    return dispatch(h, [&](auto& s) {
        for (int i = 0; i < n; ++i) s.step();
    });
}

// Sub-steps by id: 0 reset, 1 p2g, 2 gravity, 3 update velocities(dt),
// 4 grid collision(dt), 5 g2p(dt, blend), 6 particles collision(dt),
// 7 advect(dt), 8 initialize.
__attribute__((visibility("default"))) int orc_substep(void* h, int which, float dt, float blend) {
    return dispatch(h, [&](auto& s) {
        using R = std::remove_reference_t<decltype(s.h)>;
        switch (which) {
            case 0: s.reset_grid(); break;
            case 1: s.particles_to_grid(); break;
            case 2: s.compute_gravity_forces(); break;
            case 3: s.update_grid_velocities(R(dt)); break;
            case 4: s.grid_collision(R(dt)); break;
            case 5: s.grid_to_particles(R(dt), R(blend)); break;
            case 6: s.particles_collision(R(dt)); break;
            case 7: s.advect_particles(R(dt)); break;
            case 8: s.initialize(); break;
            default: break;
        }
    });
}

__attribute__((visibility("default"))) int orc_read_particles(void* h, mpm_particle_state* out) {
    if (!out) return -1;
    return dispatch(h, [&](auto& s) { to_state(s, out); });
}

__attribute__((visibility("default"))) int orc_read_particles64(void* h, double* pos, double* vel, double* Fe,
                                                                double* Fp, double* B) {
    return dispatch(h, [&](auto& s) { to_state64(s, pos, vel, Fe, Fp, B); });
}

__attribute__((visibility("default"))) int orc_read_grid(void* h, mpm_grid_host* g) {
    if (!g) return -1;
    return dispatch(h, [&](auto& s) { to_grid(s, g); });
}

__attribute__((visibility("default"))) int orc_set_particle_collision(void* h, int on) {
    return dispatch(h, [&](auto& s) { s.particle_collision = on != 0; });
}

// Wall-clock seconds for n steps (the CPU baseline timer).
__attribute__((visibility("default"))) double orc_time_steps(void* h, int n) {
    auto t0 = std::chrono::steady_clock::now();
    orc_step(h, n);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ---- standalone pieces for unit tests
__attribute__((visibility("default"))) void orc_svd3(const double* A, int64_t n, double* U, double* s, double* V,
                                                     int precision) {
    for (int64_t i = 0; i < n; ++i) {
        if (precision == 64) {
            double a[3][3], u[3][3], sv[3], v[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) a[r][c] = A[9 * i + 3 * r + c];
            svd3_f64(a, u, sv, v);
            for (int r = 0; r < 3; ++r) {
                s[3 * i + r] = sv[r];
                for (int c = 0; c < 3; ++c) { U[9 * i + 3 * r + c] = u[r][c]; V[9 * i + 3 * r + c] = v[r][c]; }
            }
        } else {
            M3<float> a, u, v;
            V3<float> sv;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) a[r][c] = float(A[9 * i + 3 * r + c]);
            svd3(a, u, sv, v);
            for (int r = 0; r < 3; ++r) {
                s[3 * i + r] = sv[r];
                for (int c = 0; c < 3; ++c) { U[9 * i + 3 * r + c] = u[r][c]; V[9 * i + 3 * r + c] = v[r][c]; }
            }
        }
    }
}

// N(x) and N'(x) of one interpolation type, FP32 restatement (precision 32)
// or FP64 (precision 64).
__attribute__((visibility("default"))) void orc_weights(int type, const double* x, int64_t n, double* N, double* Nd,
                                                        int precision) {
    if (precision == 64) {
        Interpolator<double> I(type);
        for (int64_t i = 0; i < n; ++i) { N[i] = I.N(x[i]); Nd[i] = I.Nd(x[i]); }
    } else {
        Interpolator<float> I(type);
        for (int64_t i = 0; i < n; ++i) { N[i] = I.N(float(x[i])); Nd[i] = I.Nd(float(x[i])); }
    }
}

// Stress V0-free P F^T of one particle state (model as given), FP64 or FP32.
__attribute__((visibility("default"))) void orc_stress(const double* Fe, const double* Fp, double mu0, double lambda0,
                                                       double xi, int model, double* out, int precision) {
    auto run = [&](auto tag) {
        using R = decltype(tag);
        Particle<R> p;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) { p.Fe[r][c] = R(Fe[3 * r + c]); p.Fp[r][c] = R(Fp[3 * r + c]); }
        p.mu0 = R(mu0); p.lambda0 = R(lambda0); p.hardening = R(xi); p.model = model; p.volume0 = 1;
        M3<R> P = model == MPM_MODEL_NEO_HOOKEAN ? neo_hookean_PFt(p) : fixed_corotated_PFt(p);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out[3 * r + c] = double(P[r][c]);
    };
    if (precision == 64) run(double(0)); else run(float(0));
}

// updateDeformationGradient of n particles (src/material_point3d.cu:53-70):
// Fe, Fp (9n each, row-major, updated in place) under grad (9n) and dt,
// with critical compression crit_c and stretch crit_s.
__attribute__((visibility("default"))) void orc_update_F(double* Fe, double* Fp, const double* grad, int64_t n,
                                                         double dt, double crit_c, double crit_s, int precision) {
    auto run = [&](auto tag) {
        using R = decltype(tag);
        for (int64_t i = 0; i < n; ++i) {
            Particle<R> p;
            M3<R> g;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    p.Fe[r][c] = R(Fe[9 * i + 3 * r + c]);
                    p.Fp[r][c] = R(Fp[9 * i + 3 * r + c]);
                    g[r][c] = R(grad[9 * i + 3 * r + c]);
                }
            p.crit_c = R(crit_c);
            p.crit_s = R(crit_s);
            update_deformation_gradient(p, g, R(dt));
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    Fe[9 * i + 3 * r + c] = double(p.Fe[r][c]);
                    Fp[9 * i + 3 * r + c] = double(p.Fp[r][c]);
                }
        }
    };
    if (precision == 64) run(double(0)); else run(float(0));
}

// Boundary collision of one velocity (FP32 restatement).
__attribute__((visibility("default"))) void orc_collide(const float* pos, const float* vel, const float* g000,
                                                        const float* g111, float mu, float* out) {
    V3<float> r = apply_boundary_collision(v3<float>(pos[0], pos[1], pos[2]), v3<float>(vel[0], vel[1], vel[2]),
                                           v3<float>(g000[0], g000[1], g000[2]), v3<float>(g111[0], g111[1], g111[2]), mu);
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

// Density raster of n particles (writeToOpenVDB's, FP32 or FP64): dense
// out[rx*ry*rz] in node order; active[rx*ry*rz] (may be null) flags the
// nodes the reference writes.
__attribute__((visibility("default"))) void orc_density(const mpm_desc* d, const float* pos, const float* mass,
                                                        int64_t n, double* out, uint8_t* active, int precision) {
    std::vector<uint8_t> act;
    auto run = [&](auto tag) {
        using R = decltype(tag);
        std::vector<R> v;
        density_raster<R>(*d, pos, mass, n, v, active ? &act : nullptr);
        for (size_t i = 0; i < v.size(); ++i) out[i] = double(v[i]);
    };
    if (precision == 64) run(double(0)); else run(float(0));
    if (active) std::memcpy(active, act.data(), act.size());
}

// Wall-clock seconds of one FP32 raster (the CPU baseline of the density
// pass: the reference rasterises on the host).
__attribute__((visibility("default"))) double orc_time_density(const mpm_desc* d, const float* pos,
                                                               const float* mass, int64_t n) {
    std::vector<float> v;
    auto t0 = std::chrono::steady_clock::now();
    density_raster<float>(*d, pos, mass, n, v, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// Cell keys for n positions under a descriptor.
__attribute__((visibility("default"))) void orc_keys(const mpm_desc* d, const float* pos, int64_t n, uint32_t* keys) {
    KeyParams k = key_params(*d);
    for (int64_t i = 0; i < n; ++i) keys[i] = cell_key(k, pos + 3 * i);
}

// Stable argsort of keys (std::stable_sort): order[i] = index of the i-th
// smallest key, ties in input order.
__attribute__((visibility("default"))) void orc_stable_argsort(const uint32_t* keys, int64_t n, uint32_t* order) {
    std::vector<uint32_t> idx(static_cast<size_t>(n));
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::memcpy(order, idx.data(), size_t(n) * sizeof(uint32_t));
}

}  // extern "C"
