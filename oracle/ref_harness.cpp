// ref_harness.cpp - drives the reference's own hot-path kernels, compiled from
// the unmodified reference sources (oracle/Makefile target `ref`), to produce
// golden vectors for pinning the oracle.  TEST INFRASTRUCTURE ONLY; built into
// oracle/_ref/ (never committed, never linked into the product).
//
// Reference code exercised:
//   VelocitySampling::sample_neutral_velocity   (VelocitySampling.cpp:9-35)
//   CollisionKernels::hss_collision             (CollisionKernels.cpp:139-169)
//   CollisionKernels::ehss_collision            (CollisionKernels.cpp:13-137)
//   RK4Strategy::step_batch                     (RK4Strategy.cpp:172-361)
//   BorisStrategy::step                         (BorisStrategy.cpp:18-98)
//   PhysicsRng                                  (CollisionTypes.cpp:8-20)
//   Waveform / ArbitraryWaveform::evaluate      (WaveformConfig.h, WaveformConfig.cpp)
// The only harness-side physics is the test force below (the same pattern as
// the reference's own test forces, tests/integrator/test_boris_strategy.cpp:28-49):
// the analytic quadrupole field of AnalyticalFieldModel.cpp:219-248, written
// out because that translation unit needs spdlog, which this image lacks.
//
// Protocol: stdin holds whitespace-separated numbers, the first token naming
// the mode; stdout gets the results with 17 significant digits.
#include <array>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "core/config/types/DomainConfig.h"
#include "core/config/types/WaveformConfig.h"
#include "core/integrator/strategies/BorisStrategy.h"
#include "core/integrator/strategies/RK4Strategy.h"
#include "core/physics/collisions/core/CollisionGeometry.h"
#include "core/physics/collisions/core/CollisionKernels.h"
#include "core/physics/collisions/core/VelocitySampling.h"
#include "core/physics/forces/ForceContext.h"
#include "core/physics/forces/ForceRegistry.h"
#include "core/physics/forces/IForce.h"
#include "core/types/CollisionTypes.h"
#include "core/types/IonEnsemble.h"

using namespace ICARION;

namespace {

void put(double v) { std::printf("%.17g\n", v); }
void put3(const Vec3& v) { put(v.x); put(v.y); put(v.z); }

Vec3 read3() {
    Vec3 v;
    std::cin >> v.x >> v.y >> v.z;
    return v;
}

class QuadrupoleTestForce : public physics::IForce {
public:
    QuadrupoleTestForce(double r0, double V, double f, double U, double axial_V, double L)
        : r0_(r0), V_(V), f_(f), U_(U), axial_V_(axial_V), L_(L) {}

    Vec3 E(const Vec3& pos, double t) const {
        Vec3 E_total{0.0, 0.0, 0.0};
        if (std::fabs(V_) > 1e-12) {
            const double r0_sq = r0_ * r0_;
            const double omega = 2.0 * M_PI * f_;
            const double U_eff = U_ + V_ * std::cos(omega * t);
            E_total.x = -2.0 * pos.x * U_eff / r0_sq;
            E_total.y = 2.0 * pos.y * U_eff / r0_sq;
        }
        if (std::fabs(axial_V_) > 1e-12) {
            E_total.z = axial_V_ / L_;
        }
        return E_total;
    }
    Vec3 compute(const core::IonEnsemble& ens, size_t i, double t,
                 const physics::ForceContext&) const override {
        return E(ens.get_pos(i), t) * ens.charge_data()[i];
    }
    Vec3 compute_soa(const physics::ForceState& s, double t,
                     const physics::ForceContext&) const override {
        return E(s.pos, t) * s.ion_charge_C;
    }
    std::string name() const override { return "QuadrupoleTest"; }

private:
    double r0_, V_, f_, U_, axial_V_, L_;
};

// "push": integrator(0 RK4, 1 Boris) n_ions n_steps dt t0 r0 V f U axial_V L
//         then per ion: x y z vx vy vz m q
void run_push() {
    int integrator;
    size_t n;
    long long steps;
    double dt, t0, r0, V, f, U, axV, L;
    std::cin >> integrator >> n >> steps >> dt >> t0 >> r0 >> V >> f >> U >> axV >> L;
    core::IonEnsemble ens;
    ens.resize(n);
    for (size_t i = 0; i < n; ++i) {
        Vec3 p = read3();
        Vec3 v = read3();
        double m, q;
        std::cin >> m >> q;
        ens.set_pos(i, p);
        ens.set_vel(i, v);
        ens.update_species(i, "ion", m, q, 0.0, 0.0);
        ens.mass_data()[i] = m;
        ens.charge_data()[i] = q;
        ens.active_data()[i] = 1;
        ens.born_data()[i] = 1;
    }
    config::DomainConfig dom;
    auto reg = std::make_shared<physics::ForceRegistry>(dom);
    reg->add_force(std::make_unique<QuadrupoleTestForce>(r0, V, f, U, axV, L));
    std::vector<std::shared_ptr<physics::ForceRegistry>> regs{reg};
    std::vector<int> domains(n, 0);
    integrator::RK4Strategy rk4;
    rk4.set_parallel_enabled(false);
    integrator::BorisStrategy boris;
    double t = t0;
    for (long long k = 0; k < steps; ++k) {
        if (integrator == 0) {
            rk4.step_batch(ens, t, dt, regs, domains);
        } else {
            for (size_t i = 0; i < n; ++i) boris.step(ens, i, t, dt, *reg);
        }
        t = t + dt;  // t_i += dt, engine time = max_i t_i
    }
    for (size_t i = 0; i < n; ++i) {
        put3(ens.get_pos(i));
        put3(ens.get_vel(i));
    }
}

// "neutral": seed T m fx fy fz
// waveform: type clamp interp p0..p7 n_points (t v)... n_t t...
// prints the waveform value at each t
void run_waveform() {
    int type, clamp, interp;
    double p[8];
    std::cin >> type >> clamp >> interp;
    for (double& x : p) std::cin >> x;
    size_t n_points;
    std::cin >> n_points;
    config::ArbitraryWaveform arb;
    for (size_t i = 0; i < n_points; ++i) {
        double t, v;
        std::cin >> t >> v;
        arb.times_s.push_back(t);
        arb.values.push_back(v);
    }
    arb.interp = interp == 1 ? config::ArbitraryWaveform::Interpolation::Step
               : interp == 2 ? config::ArbitraryWaveform::Interpolation::Cubic
                             : config::ArbitraryWaveform::Interpolation::Linear;
    config::Waveform w;
    switch (type) {
        case 0: w.data = config::ConstantWaveform{p[0]}; break;
        case 1: w.data = config::LinearWaveform{p[0], p[1], p[2], p[3], clamp != 0}; break;
        case 2: w.data = config::QuadraticWaveform{p[0], p[1], p[2], p[3], p[4]}; break;
        case 3: w.data = config::ExponentialWaveform{p[0], p[1], p[2], p[3], p[4], clamp != 0}; break;
        case 4: w.data = config::SinusoidalWaveform{p[0], p[1], p[2], p[3]}; break;
        case 5: w.data = config::PWMWaveform{p[0], p[1], p[2], p[3], p[4], p[5], p[6], clamp != 0}; break;
        case 6: w.data = config::PulsedWaveform{p[0], p[1], p[2], p[3]}; break;
        default: w.data = arb; break;
    }
    size_t n_t;
    std::cin >> n_t;
    for (size_t i = 0; i < n_t; ++i) {
        double t;
        std::cin >> t;
        put(w.evaluate(t));
    }
}

void run_neutral() {
    unsigned long long seed;
    double T, m;
    std::cin >> seed >> T >> m;
    Vec3 flow = read3();
    physics::PhysicsRng rng(seed);
    put3(physics::collision_core::VelocitySampling::sample_neutral_velocity(T, m, flow, rng));
}

// "hss": seed vix viy viz vnx vny vnz mi mn
void run_hss() {
    unsigned long long seed;
    std::cin >> seed;
    Vec3 vi = read3();
    Vec3 vn = read3();
    double mi, mn;
    std::cin >> mi >> mn;
    physics::PhysicsRng rng(seed);
    put3(physics::collision_core::CollisionKernels::hss_collision(vi, vn, mi, mn, rng));
}

// "ehss": seed vi(3) vn(3) mi mn ion_radius n_atoms then n_atoms x (x y z r)
void run_ehss() {
    unsigned long long seed;
    std::cin >> seed;
    Vec3 vi = read3();
    Vec3 vn = read3();
    double mi, mn, rion;
    size_t nat;
    std::cin >> mi >> mn >> rion >> nat;
    std::vector<Vec3> c(nat);
    std::vector<double> r(nat);
    for (size_t j = 0; j < nat; ++j) {
        c[j] = read3();
        std::cin >> r[j];
    }
    physics::PhysicsRng rng(seed);
    put3(physics::collision_core::CollisionKernels::ehss_collision(vi, vn, mi, mn, rion, c, r, rng));
}

// ehss_orient: CollisionKernels::ehss_collision_with_orientation with the
// rotation of CollisionGeometry::quaternion_to_rotation, sigma_eff and the
// forced-hit budget, as EHSSCollisionHandler calls it for orientation samples
void run_ehss_orient() {
    unsigned long long seed;
    std::cin >> seed;
    Vec3 vi = read3();
    Vec3 vn = read3();
    double mi, mn, rion;
    size_t nat;
    std::cin >> mi >> mn >> rion >> nat;
    std::vector<Vec3> c(nat);
    std::vector<double> r(nat);
    for (size_t j = 0; j < nat; ++j) {
        c[j] = read3();
        std::cin >> r[j];
    }
    std::array<double, 4> q;
    std::cin >> q[0] >> q[1] >> q[2] >> q[3];
    double sigma;
    std::cin >> sigma;
    double R[3][3];
    physics::collision_core::CollisionGeometry::quaternion_to_rotation(q, R);
    physics::PhysicsRng rng(seed);
    put3(physics::collision_core::CollisionKernels::ehss_collision_with_orientation(
        vi, vn, mi, mn, rion, c, r, R, rng, 256, sigma, true));
}

// quat: CollisionGeometry::quaternion_to_rotation, row-major
void run_quat() {
    std::array<double, 4> q;
    std::cin >> q[0] >> q[1] >> q[2] >> q[3];
    double R[3][3];
    physics::collision_core::CollisionGeometry::quaternion_to_rotation(q, R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) std::printf("%.17g%c", R[i][j], (i == 2 && j == 2) ? '\n' : ' ');
}

}  // namespace

int main() {
    std::string mode;
    while (std::cin >> mode) {
        if (mode == "push") run_push();
        else if (mode == "neutral") run_neutral();
        else if (mode == "hss") run_hss();
        else if (mode == "ehss") run_ehss();
        else if (mode == "waveform") run_waveform();
        else if (mode == "ehss_orient") run_ehss_orient();
        else if (mode == "quat") run_quat();
        else {
            std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
            return 2;
        }
    }
    return 0;
}
