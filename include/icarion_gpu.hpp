// icarion_gpu.hpp - C++17 host adaptor over the C ABI of icarion_gpu.h.
//
// BatchStepper follows the contract of the reference's batch hooks:
//   IIntegrationStrategy::supports_batch / step_batch
//     (src/core/integrator/strategies/IIntegrationStrategy.h:106-132)
//   ICollisionHandler::supports_batch / handle_batch
//     (src/core/physics/collisions/ICollisionHandler.h:109-134)
// step_batch() returns false instead of throwing when the device path cannot
// run (no device, unsupported configuration, HIP failure), so the caller falls
// back to its CPU loop, as GPUIntegrationStrategy does
// (GPUIntegrationStrategy.cpp:22-34, 277-279).  The fused device step covers
// push + collision + boundary elimination of one engine step, so an engine that
// uses it skips its own perform_collisions / finalize for that step.
//
// Resident mode (upload once, step many, download once) avoids the per-step
// PCIe round trip of the reference's GPU helpers (GPUIntegrationHelper.cpp:188-228).
#ifndef ICARION_GPU_HPP
#define ICARION_GPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "icarion_gpu.h"

namespace icarion {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline std::string last_error() {
    char buf[1024];
    icarion_gpu_last_error(buf, sizeof(buf));
    return std::string(buf);
}

inline void check(int rc) {
    if (rc != ICARION_OK) throw Error(rc, last_error());
}

class BatchStepper {
public:
    BatchStepper() = default;
    explicit BatchStepper(const icarion_gpu_desc& desc) { check(icarion_gpu_create(&desc, &h_)); }
    ~BatchStepper() { reset(); }
    BatchStepper(const BatchStepper&) = delete;
    BatchStepper& operator=(const BatchStepper&) = delete;
    BatchStepper(BatchStepper&& o) noexcept : h_(std::exchange(o.h_, nullptr)), ready_(o.ready_) {}
    BatchStepper& operator=(BatchStepper&& o) noexcept {
        if (this != &o) {
            reset();
            h_ = std::exchange(o.h_, nullptr);
            ready_ = o.ready_;
        }
        return *this;
    }

    // True when a device is visible and the handle exists
    // (IIntegrationStrategy::supports_batch).
    bool supports_batch() const { return h_ != nullptr && ready_; }

    // Throwing configuration calls.
    // atoms_xyzr: the EHSS hard spheres, float x, y, z, r per atom (the device's
    // atom table is float; icarion::config::species_table builds it).
    void configure(const icarion_domain_desc& domain, const icarion_physics_desc& physics,
                   const std::vector<icarion_species_desc>& species,
                   const std::vector<float>& atoms_xyzr = {}) {
        check(icarion_gpu_set_domain(h_, &domain, &physics));
        check(icarion_gpu_set_species(h_, species.data(), static_cast<uint32_t>(species.size()),
                                      atoms_xyzr.empty() ? nullptr : atoms_xyzr.data(),
                                      static_cast<uint32_t>(atoms_xyzr.size() / 4), 0));
        ready_ = true;
    }
    // A gas mixture: components, sigma per [species][component] and (EHSS) the
    // components' contact radii (icarion_gpu_set_gas_mixture / _set_gas_radii).
    void set_gas_mixture(const std::vector<icarion_gas_component>& components, const std::vector<double>& sigma_m2,
                         uint32_t n_species, const std::vector<double>& radius_m = {}) {
        check(icarion_gpu_set_gas_mixture(h_, components.data(), static_cast<uint32_t>(components.size()),
                                          sigma_m2.data(), n_species));
        if (!radius_m.empty())
            check(icarion_gpu_set_gas_radii(h_, radius_m.data(), static_cast<uint32_t>(radius_m.size())));
    }
    // The gas flow profile (icarion_gpu_set_gas_flow): ICARION_FLOW_CONSTANT
    // clears, ICARION_FLOW_AXIAL_PARABOLIC with its centre-line speed.
    void set_gas_flow(int32_t flow_model, double axial_max_velocity_m_s) {
        check(icarion_gpu_set_gas_flow(h_, flow_model, axial_max_velocity_m_s));
    }
    // Several domains (icarion_gpu_set_domains): one descriptor and end aperture
    // per DomainConfig, in the config's order.
    // domain_species: empty, or per domain its species rows (the same species
    // in the same order, cross sections / CCS / K0 resolved for that domain's
    // gas); an empty entry leaves that domain on `species`.
    void configure_domains(const std::vector<icarion_domain_desc>& domains, const std::vector<double>& end_aperture_m,
                           const icarion_physics_desc& physics, const std::vector<icarion_species_desc>& species,
                           const std::vector<std::vector<icarion_species_desc>>& domain_species = {}) {
        if (!end_aperture_m.empty() && end_aperture_m.size() != domains.size())
            throw Error(ICARION_ERR_INVALID, "one end aperture per domain");
        if (!domain_species.empty() && domain_species.size() != domains.size())
            throw Error(ICARION_ERR_INVALID, "one species table (or none) per domain");
        check(icarion_gpu_set_domains(h_, domains.data(), end_aperture_m.empty() ? nullptr : end_aperture_m.data(),
                                      static_cast<uint32_t>(domains.size()), &physics));
        check(icarion_gpu_set_species(h_, species.data(), static_cast<uint32_t>(species.size()), nullptr, 0, 0));
        for (size_t k = 0; k < domain_species.size(); ++k)
            if (!domain_species[k].empty())
                check(icarion_gpu_set_domain_species(h_, static_cast<uint32_t>(k), domain_species[k].data(),
                                                     static_cast<uint32_t>(domain_species[k].size())));
        ready_ = true;
    }
    // Step control of ICARION_INTEGRATOR_RK45 (RK45Strategy(const AdaptiveConfig&)).
    void set_rk45(const icarion_rk45_desc& cfg) { check(icarion_gpu_set_rk45(h_, &cfg)); }
    void upload_field_array(uint32_t nx, uint32_t ny, uint32_t nz, const double origin[3],
                            const double spacing[3], const float* E_xyz) {
        check(icarion_gpu_upload_field_array(h_, nx, ny, nz, origin[0], spacing[0], origin[1],
                                             spacing[1], origin[2], spacing[2], E_xyz, 0));
    }

    // Resident mode.
    void upload(const icarion_ion_soa& ions) { check(icarion_gpu_upload_state(h_, &ions)); }
    icarion_step_stats advance(double t0, double dt, uint64_t n_steps, uint64_t step_offset) {
        icarion_step_stats st{};
        check(icarion_gpu_step(h_, t0, dt, n_steps, step_offset, &st));
        return st;
    }
    void download(icarion_ion_soa& ions) { check(icarion_gpu_download_state(h_, &ions)); }
    // Resident stepping with OutputManager's trajectory sampling: fn receives each
    // snapshot; next_write_s carries the schedule from one call to the next
    // (initially the write interval, as OutputManager::initialize sets it).
    icarion_step_stats advance_sampled(double t0, double dt, uint64_t n_steps, uint64_t step_offset,
                                       double write_interval_s, bool write_first, double& next_write_s,
                                       icarion_snapshot_fn fn, void* user) {
        icarion_step_stats st{};
        check(icarion_gpu_step_snapshots(h_, t0, dt, n_steps, step_offset, write_interval_s,
                                         write_first ? 1 : 0, &next_write_s, fn, user, &st));
        return st;
    }

    // One engine step on a host-owned ensemble.  false = "not handled, run the
    // CPU path" (the reference's step_batch / handle_batch return value).
    bool step_batch(icarion_ion_soa& ions, double t, double dt, uint64_t step_index,
                    icarion_step_stats* stats = nullptr) noexcept {
        if (!supports_batch()) return false;
        icarion_step_stats st{};
        const int rc = icarion_gpu_step_batch(h_, &ions, t, dt, step_index, &st);
        if (rc != ICARION_OK) return false;
        if (stats) *stats = st;
        return true;
    }

    icarion_gpu* handle() const { return h_; }

    // Factory that never throws: an empty stepper when no device is visible or
    // creation fails (caller keeps using its CPU strategy).
    static BatchStepper try_create(const icarion_gpu_desc& desc) {
        BatchStepper s;
        if (icarion_gpu_device_count() <= 0) return s;
        if (icarion_gpu_create(&desc, &s.h_) != ICARION_OK) s.h_ = nullptr;
        return s;
    }

private:
    void reset() {
        if (h_) icarion_gpu_destroy(h_);
        h_ = nullptr;
        ready_ = false;
    }
    icarion_gpu* h_ = nullptr;
    bool ready_ = false;
};

// Many independent single-domain clouds in one persistent launch
// (icarion_gpu_batch_*).  Each cloud's result is that of a BatchStepper of its
// own run in resident mode with the same descriptors.  The clouds' host SoA
// arrays must stay valid from upload() to download().
class CloudBatch {
public:
    explicit CloudBatch(const icarion_gpu_desc& desc) { check(icarion_gpu_create(&desc, &h_)); }
    ~CloudBatch() {
        if (h_) icarion_gpu_destroy(h_);
    }
    CloudBatch(const CloudBatch&) = delete;
    CloudBatch& operator=(const CloudBatch&) = delete;

    void upload(std::vector<icarion_cloud_desc> clouds) {
        check(icarion_gpu_batch_upload(h_, clouds.data(), static_cast<uint32_t>(clouds.size())));
        clouds_ = std::move(clouds);
    }
    // Picks the launch shape per precision class on the resident batch (its
    // state is restored); returns the FP64 and FP32 choices.
    std::pair<icarion_batch_tuning, icarion_batch_tuning> autotune(uint32_t trial_steps = 300) {
        icarion_batch_tuning b64{}, b32{};
        check(icarion_gpu_batch_autotune(h_, trial_steps, &b64, &b32, nullptr, nullptr));
        return {b64, b32};
    }
    // One graph launch over every cloud; per-cloud counters and completion times.
    std::vector<icarion_cloud_stats> run(double* device_ms = nullptr) {
        std::vector<icarion_cloud_stats> st(clouds_.size());
        check(icarion_gpu_batch_run(h_, nullptr, nullptr, st.data(), device_ms));
        return st;
    }
    void download() {
        check(icarion_gpu_batch_download(h_, clouds_.data(), static_cast<uint32_t>(clouds_.size())));
    }
    icarion_gpu* handle() const { return h_; }

private:
    icarion_gpu* h_ = nullptr;
    std::vector<icarion_cloud_desc> clouds_;
};

}  // namespace icarion

#endif  // ICARION_GPU_HPP
