// mpm_solver3d.cpp -- facade classes over the C-ABI (include/mpm_hip.h).
#include "mpm_solver3d.h"

#include <stdexcept>
#include <string>

#include "apic_mpm_solver3d.h"
#include "mpm_hip.h"

namespace chains {

namespace {
void check(int rc, const char* what) {
    if (rc != MPM_OK) throw std::runtime_error(std::string(what) + ": " + mpm_last_error());
}

mpm_particle_state pack(const std::vector<MaterialPoint3D>& P, const std::vector<Matrix3f>* B,
                        std::vector<float>& f, std::vector<int32_t>& model) {
    const size_t n = P.size();
    // layout: volume0 n, mass n, pos 3n, vel 3n, Fe 9n, Fp 9n, mu0 n, lam0 n, xi n, thc n, ths n, B 9n
    f.assign(n * 40, 0.0f);
    model.assign(n, 0);
    float* volume0 = f.data();
    float* mass = volume0 + n;
    float* pos = mass + n;
    float* vel = pos + 3 * n;
    float* Fe = vel + 3 * n;
    float* Fp = Fe + 9 * n;
    float* mu0 = Fp + 9 * n;
    float* lam0 = mu0 + n;
    float* xi = lam0 + n;
    float* thc = xi + n;
    float* ths = thc + n;
    float* Bp = ths + n;
    for (size_t i = 0; i < n; ++i) {
        const MaterialPoint3D& p = P[i];
        volume0[i] = p.volume0;
        mass[i] = p.mass;
        for (int a = 0; a < 3; ++a) {
            pos[3 * i + a] = p.position(a);
            vel[3 * i + a] = p.velocity(a);
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                Fe[9 * i + 3 * r + c] = p.deformation_gradient_elastic(r, c);
                Fp[9 * i + 3 * r + c] = p.deformation_gradient_plastic(r, c);
                if (B) Bp[9 * i + 3 * r + c] = (*B)[i](r, c);
            }
        mu0[i] = p.mu0;
        lam0[i] = p.lambda0;
        xi[i] = p.hardening_coefficient;
        thc[i] = p.critical_compression;
        ths[i] = p.critical_stretch;
        model[i] = (int32_t)p.constitutive_model;
    }
    mpm_particle_state s{};
    s.count = (int64_t)n;
    s.volume0 = volume0; s.mass = mass; s.position = pos; s.velocity = vel; s.Fe = Fe; s.Fp = Fp;
    s.mu0 = mu0; s.lambda0 = lam0; s.hardening = xi; s.critical_compression = thc; s.critical_stretch = ths;
    s.model = model.data();
    s.B = B ? Bp : nullptr;
    return s;
}
}  // namespace

MPMSolverHandle::MPMSolverHandle(const std::vector<MaterialPoint3D>& particles, const std::vector<Matrix3f>* B,
                                 Vector3f gridOrigin, Vector3i gridResolution, float gridStride,
                                 float gridBoundaryFrictionCoefficient, float blendCoefficient,
                                 InterpolationType interpolationType, float deltaTime, int transfer, float gravity)
    : _n((int)particles.size()), _blend_coefficient(blendCoefficient), _delta_time(deltaTime) {
    mpm_desc d{};
    for (int i = 0; i < 3; ++i) {
        d.origin[i] = gridOrigin(i);
        d.resolution[i] = gridResolution(i);
    }
    d.stride = gridStride;
    d.friction = gridBoundaryFrictionCoefficient;
    d.blend = blendCoefficient;
    d.interp = (int32_t)interpolationType;
    d.transfer = transfer;
    d.dt = deltaTime;
    d.gravity = gravity;
    d.particle_collision = 0;
    std::vector<float> f;
    std::vector<int32_t> model;
    mpm_particle_state s = pack(particles, B, f, model);
    // The constructor runs initialize() (reference src/mpm_solver3d.cu:88-90).
    check(mpm_create_state(&d, &s, 1, &_h), "MPM solver construction");
}

MPMSolverHandle::MPMSolverHandle(MPMSolverHandle&& o) noexcept
    : _enable_particles_collision(o._enable_particles_collision), _h(o._h), _n(o._n),
      _blend_coefficient(o._blend_coefficient), _delta_time(o._delta_time) {
    o._h = nullptr;
}

MPMSolverHandle& MPMSolverHandle::operator=(MPMSolverHandle&& o) noexcept {
    if (this != &o) {
        if (_h) mpm_destroy(_h);
        _enable_particles_collision = o._enable_particles_collision;
        _h = o._h;
        _n = o._n;
        _blend_coefficient = o._blend_coefficient;
        _delta_time = o._delta_time;
        o._h = nullptr;
    }
    return *this;
}

MPMSolverHandle::~MPMSolverHandle() {
    if (_h) mpm_destroy(_h);
}

void MPMSolverHandle::syncCollisionFlag() {
    check(mpm_set_particles_collision(_h, _enable_particles_collision ? 1 : 0), "particle collision flag");
}

void MPMSolverHandle::simulateOneStep() {
    syncCollisionFlag();
    check(mpm_step(_h, 1), "simulateOneStep");
}

void MPMSolverHandle::switchIfEnableParticlesCollision() { _enable_particles_collision = !_enable_particles_collision; }
void MPMSolverHandle::initialize() { check(mpm_initialize(_h), "initialize"); }
void MPMSolverHandle::resetGrid() { check(mpm_reset_grid(_h), "resetGrid"); }
void MPMSolverHandle::particlesToGrid() { check(mpm_particles_to_grid(_h), "particlesToGrid"); }
void MPMSolverHandle::computeExternalForces() { check(mpm_compute_external_forces(_h), "computeExternalForces"); }
void MPMSolverHandle::updateGridVelocities(float dt) { check(mpm_update_grid_velocities(_h, dt), "updateGridVelocities"); }
void MPMSolverHandle::solveLinearSystem(float dt) { check(mpm_solve_linear_system(_h, dt), "solveLinearSystem"); }
void MPMSolverHandle::advectParticles(float dt) { check(mpm_advect_particles(_h, dt), "advectParticles"); }
void MPMSolverHandle::computeGravityForces() { check(mpm_compute_gravity_forces(_h), "computeGravityForces"); }
void MPMSolverHandle::gridCollision(float dt) { check(mpm_grid_collision(_h, dt), "gridCollision"); }
void MPMSolverHandle::particlesCollision(float dt) { check(mpm_particles_collision(_h, dt), "particlesCollision"); }
void MPMSolverHandle::copyPositionsToHost(float* xyz1) const {
    check(mpm_read_positions4(_h, xyz1), "copyPositionsToHost");
}
void MPMSolverHandle::writeToOpenVDB(std::string filePath) {
    check(mpm_write_density(_h, filePath.c_str()), "writeToOpenVDB");
}
void MPMSolverHandle::synchronize() const { check(mpm_synchronize(_h), "synchronize"); }

void MPMSolverHandle::readState(std::vector<MaterialPoint3D>& out, std::vector<Matrix3f>* B) const {
    const size_t n = (size_t)_n;
    std::vector<float> f(n * 40);
    std::vector<int32_t> model(n);
    float* base = f.data();
    mpm_particle_state s{};
    s.count = (int64_t)n;
    s.volume0 = base; s.mass = base + n; s.position = base + 2 * n; s.velocity = base + 5 * n; s.Fe = base + 8 * n;
    s.Fp = base + 17 * n; s.mu0 = base + 26 * n; s.lambda0 = base + 27 * n; s.hardening = base + 28 * n;
    s.critical_compression = base + 29 * n; s.critical_stretch = base + 30 * n; s.model = model.data();
    s.B = B ? base + 31 * n : nullptr;
    check(mpm_read_particles(_h, &s), "readParticles");
    out.resize(n);
    if (B) B->resize(n);
    for (size_t i = 0; i < n; ++i) {
        MaterialPoint3D& p = out[i];
        p.volume0 = s.volume0[i];
        p.mass = s.mass[i];
        for (int a = 0; a < 3; ++a) {
            p.position(a) = s.position[3 * i + a];
            p.velocity(a) = s.velocity[3 * i + a];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                p.deformation_gradient_elastic(r, c) = s.Fe[9 * i + 3 * r + c];
                p.deformation_gradient_plastic(r, c) = s.Fp[9 * i + 3 * r + c];
                if (B) (*B)[i](r, c) = s.B[9 * i + 3 * r + c];
            }
        p.mu0 = s.mu0[i];
        p.lambda0 = s.lambda0[i];
        p.hardening_coefficient = s.hardening[i];
        p.critical_compression = s.critical_compression[i];
        p.critical_stretch = s.critical_stretch[i];
        p.constitutive_model = (ConstitutiveModel)s.model[i];
    }
}

MPMSolver3D::MPMSolver3D(const std::vector<MaterialPoint3D>& particles, Vector3f gridOrigin, Vector3i gridResolution,
                         float gridStride, float gridBoundaryFrictionCoefficient, float blendCoefficient,
                         InterpolationType interpolationType, float deltaTime, float gravity)
    : MPMSolverHandle(particles, nullptr, gridOrigin, gridResolution, gridStride, gridBoundaryFrictionCoefficient,
                      blendCoefficient, interpolationType, deltaTime, MPM_TRANSFER_PIC_FLIP, gravity) {}

void MPMSolver3D::gridToParticles(float dt, float blend) {
    check(mpm_grid_to_particles(_h, dt, blend), "gridToParticles");
}

static std::vector<MaterialPoint3D> base_points(const std::vector<APICMaterialPoint3D>& P) {
    return std::vector<MaterialPoint3D>(P.begin(), P.end());
}
static std::vector<Matrix3f> b_matrices(const std::vector<APICMaterialPoint3D>& P) {
    std::vector<Matrix3f> B(P.size());
    for (size_t i = 0; i < P.size(); ++i) B[i] = P[i].B;
    return B;
}

APICMPMSolver3D::APICMPMSolver3D(const std::vector<APICMaterialPoint3D>& particles, Vector3f gridOrigin,
                                 Vector3i gridResolution, float gridStride, float gridBoundaryFrictionCoefficient,
                                 InterpolationType interpolationType, float deltaTime, float gravity)
    : MPMSolverHandle(base_points(particles), nullptr, gridOrigin, gridResolution, gridStride,
                      gridBoundaryFrictionCoefficient, 0.0f, interpolationType, deltaTime, MPM_TRANSFER_APIC,
                      gravity) {
    // B of freshly constructed particles is zero; a non-zero B is restored.
    std::vector<Matrix3f> B = b_matrices(particles);
    bool nonzero = false;
    for (auto& m : B)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) nonzero |= m(r, c) != 0.0f;
    if (nonzero) {
        std::vector<MaterialPoint3D> cur;
        readState(cur, nullptr);
        std::vector<float> f;
        std::vector<int32_t> model;
        mpm_particle_state s = pack(cur, &B, f, model);
        check(mpm_write_particles(_h, &s), "APIC B restore");
    }
}

void APICMPMSolver3D::gridToParticles(float dt) {
    check(mpm_grid_to_particles(_h, dt, 0.0f), "gridToParticles");
}

void APICMPMSolver3D::readParticles(std::vector<APICMaterialPoint3D>& out) const {
    std::vector<MaterialPoint3D> base;
    std::vector<Matrix3f> B;
    readState(base, &B);
    out.resize(base.size());
    for (size_t i = 0; i < base.size(); ++i) {
        static_cast<MaterialPoint3D&>(out[i]) = base[i];
        out[i].B = B[i];
    }
}

}  // namespace chains
