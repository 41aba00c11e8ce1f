// trajectory_h5.cpp - the /ions and /trajectory groups of the reference's
// simulation output file (include/icarion_h5.h), on the HDF5 C library.
//
// The snapshots come from icarion_gpu_step_snapshots (or any caller) one at a
// time; they are held in host buffers and appended in batches, the way
// OutputManager buffers log_step rows before
// HDF5Writer::append_trajectory_batch_flat (src/core/io/hdf5Writer.cpp:530-709)
// writes them: every dataset extendible along T, chunked (min(T0, 100), N[, 3])
// (time: min(T0, 1000)) and deflate-compressed.
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/icarion_h5.h"

namespace {

thread_local std::string g_error;

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ICARION_OK;
    } catch (const Failure& e) {
        g_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return ICARION_ERR_INVALID;
    } catch (const std::exception& e) {
        g_error = e.what();
        return ICARION_ERR_INVALID;
    }
}

hid_t check_id(hid_t id, const char* what) {
    if (id < 0) fail(ICARION_ERR_INVALID, std::string("HDF5: ") + what + " failed");
    return id;
}

void check_rc(herr_t rc, const char* what) {
    if (rc < 0) fail(ICARION_ERR_INVALID, std::string("HDF5: ") + what + " failed");
}

// Closes an HDF5 object id at scope exit.
struct Id {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Id() {
        if (id >= 0 && close) close(id);
    }
    Id(const Id&) = delete;
    Id& operator=(const Id&) = delete;
    operator hid_t() const { return id; }
};

hid_t varlen_string_type() {
    const hid_t t = check_id(H5Tcopy(H5T_C_S1), "H5Tcopy");
    if (H5Tset_size(t, H5T_VARIABLE) < 0) {
        H5Tclose(t);
        fail(ICARION_ERR_INVALID, "HDF5: H5Tset_size failed");
    }
    return t;
}

void write_1d(hid_t group, const char* name, hid_t type, const void* data, hsize_t n) {
    Id space(check_id(H5Screate_simple(1, &n, nullptr), "H5Screate_simple"), H5Sclose);
    Id ds(check_id(H5Dcreate2(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), name), H5Dclose);
    if (n) check_rc(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), name);
}

}  // namespace

struct icarion_h5 {
    hid_t file = -1;
    hid_t traj = -1;
    hid_t ions = -1;
    size_t n = 0;
    int deflate = 2;
    size_t batch = 100;
    std::vector<std::string> pool;
    // buffered rows
    size_t rows = 0;
    uint64_t written = 0;
    int per_ion_time = -1;   // unknown until the first row: 0 no, 1 yes
    std::vector<double> t, pos, vel, tpi;
    std::vector<int32_t> dom;
    std::vector<uint32_t> spec;
};

namespace {

// Append rows [T, dims...] to an extendible dataset, creating it with the
// first batch (chunk along T = min(first batch, t_chunk)).
void append_rows(icarion_h5* h, const char* name, hid_t type, const void* data, hsize_t rows, int rank,
                 const hsize_t* inner, hsize_t t_chunk) {
    hsize_t dims[3] = {rows, rank > 1 ? inner[0] : 0, rank > 2 ? inner[1] : 0};
    if (H5Lexists(h->traj, name, H5P_DEFAULT) > 0) {
        Id ds(check_id(H5Dopen2(h->traj, name, H5P_DEFAULT), name), H5Dclose);
        Id fs0(check_id(H5Dget_space(ds), "H5Dget_space"), H5Sclose);
        hsize_t old[3] = {0, 0, 0};
        check_rc(H5Sget_simple_extent_dims(fs0, old, nullptr) == rank ? 0 : -1, "extent rank");
        hsize_t ext[3] = {old[0] + rows, old[1], old[2]};
        check_rc(H5Dset_extent(ds, ext), "H5Dset_extent");
        Id fs(check_id(H5Dget_space(ds), "H5Dget_space"), H5Sclose);
        hsize_t start[3] = {old[0], 0, 0};
        check_rc(H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, dims, nullptr), "H5Sselect_hyperslab");
        Id ms(check_id(H5Screate_simple(rank, dims, nullptr), "H5Screate_simple"), H5Sclose);
        check_rc(H5Dwrite(ds, type, ms, fs, H5P_DEFAULT, data), name);
        return;
    }
    hsize_t maxd[3] = {H5S_UNLIMITED, dims[1], dims[2]};
    Id space(check_id(H5Screate_simple(rank, dims, maxd), "H5Screate_simple"), H5Sclose);
    Id plist(check_id(H5Pcreate(H5P_DATASET_CREATE), "H5Pcreate"), H5Pclose);
    hsize_t chunk[3] = {std::min<hsize_t>(std::max<hsize_t>(rows, 1), t_chunk), dims[1], dims[2]};
    check_rc(H5Pset_chunk(plist, rank, chunk), "H5Pset_chunk");
    if (h->deflate > 0) check_rc(H5Pset_deflate(plist, static_cast<unsigned>(h->deflate)), "H5Pset_deflate");
    Id ds(check_id(H5Dcreate2(h->traj, name, type, space, H5P_DEFAULT, plist, H5P_DEFAULT), name), H5Dclose);
    check_rc(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), name);
}

void flush_rows(icarion_h5* h) {
    if (h->rows == 0) return;
    const hsize_t T = h->rows;
    const hsize_t nv[2] = {h->n, 3};
    const hsize_t n1[1] = {h->n};
    append_rows(h, "time", H5T_NATIVE_DOUBLE, h->t.data(), T, 1, nullptr, 1000);
    append_rows(h, "positions", H5T_NATIVE_DOUBLE, h->pos.data(), T, 3, nv, 100);
    append_rows(h, "velocities", H5T_NATIVE_DOUBLE, h->vel.data(), T, 3, nv, 100);
    append_rows(h, "domain_indices", H5T_NATIVE_INT32, h->dom.data(), T, 2, n1, 100);
    append_rows(h, "species_id_indices", H5T_NATIVE_UINT32, h->spec.data(), T, 2, n1, 100);
    if (h->per_ion_time == 1) append_rows(h, "time_per_ion", H5T_NATIVE_DOUBLE, h->tpi.data(), T, 2, n1, 100);
    if (!h->pool.empty()) {
        // compatibility dataset: the species name of every ion (hdf5Writer.cpp:662-700)
        std::vector<const char*> names(T * h->n);
        for (size_t k = 0; k < names.size(); ++k) {
            const uint32_t s = h->spec[k];
            names[k] = s < h->pool.size() ? h->pool[s].c_str() : "";
        }
        Id st(varlen_string_type(), H5Tclose);
        append_rows(h, "species_ids", st, names.data(), T, 2, n1, 100);
    }
    check_rc(H5Fflush(h->file, H5F_SCOPE_GLOBAL), "H5Fflush");
    h->written += T;
    h->rows = 0;
}

void close_all(icarion_h5* h) {
    if (h->traj >= 0) H5Gclose(h->traj);
    if (h->ions >= 0) H5Gclose(h->ions);
    if (h->file >= 0) H5Fclose(h->file);
    h->traj = h->ions = h->file = -1;
}

}  // namespace

extern "C" {

size_t icarion_h5_last_error(char* buf, size_t len) {
    if (buf && len) {
        const size_t k = std::min(len - 1, g_error.size());
        std::memcpy(buf, g_error.data(), k);
        buf[k] = '\0';
    }
    return g_error.size();
}

int icarion_h5_create(const char* path, const icarion_ion_soa* ini, const char* const* species_pool,
                      uint32_t n_species_pool, int deflate_level, uint32_t batch, icarion_h5** out) {
    return guarded([&] {
        if (!out) fail(ICARION_ERR_INVALID, "out is NULL");
        *out = nullptr;
        if (!path || !ini) fail(ICARION_ERR_INVALID, "NULL argument");
        if (deflate_level < 0 || deflate_level > 9) fail(ICARION_ERR_INVALID, "deflate_level must be 0-9");
        const size_t n = ini->n;
        if (n && (!ini->x || !ini->y || !ini->z || !ini->vx || !ini->vy || !ini->vz || !ini->charge_C ||
                  !ini->species))
            fail(ICARION_ERR_INVALID, "the initial state needs positions, velocities, charges and species");
        if (species_pool && n_species_pool == 0) fail(ICARION_ERR_INVALID, "empty species pool");
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        auto h = new icarion_h5();
        try {
            h->n = n;
            h->deflate = deflate_level;
            // at most ~1 GiB of buffered rows (64 B per ion and row)
            const size_t want = batch ? batch : 100;
            const size_t cap = n ? std::max<size_t>(1, (size_t(1) << 30) / (64 * n)) : want;
            h->batch = std::max<size_t>(1, std::min(want, cap));
            if (species_pool)
                for (uint32_t s = 0; s < n_species_pool; ++s) h->pool.emplace_back(species_pool[s] ? species_pool[s] : "");
            h->file = check_id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate");
            h->ions = check_id(H5Gcreate2(h->file, "/ions", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "/ions");
            // HDF5Writer::write_ion_metadata (hdf5Writer.cpp:1656-1730)
            std::vector<std::string> names(n);
            std::vector<const char*> ptrs(n);
            for (size_t i = 0; i < n; ++i) {
                const uint32_t s = ini->species[i];
                names[i] = h->pool.empty() ? std::to_string(s) : (s < h->pool.size() ? h->pool[s] : std::string());
                ptrs[i] = names[i].c_str();
            }
            {
                Id st(varlen_string_type(), H5Tclose);
                write_1d(h->ions, "initial_species_id", st, ptrs.data(), n);
            }
            write_1d(h->ions, "initial_pos_x", H5T_NATIVE_DOUBLE, ini->x, n);
            write_1d(h->ions, "initial_pos_y", H5T_NATIVE_DOUBLE, ini->y, n);
            write_1d(h->ions, "initial_pos_z", H5T_NATIVE_DOUBLE, ini->z, n);
            write_1d(h->ions, "initial_vel_x", H5T_NATIVE_DOUBLE, ini->vx, n);
            write_1d(h->ions, "initial_vel_y", H5T_NATIVE_DOUBLE, ini->vy, n);
            write_1d(h->ions, "initial_vel_z", H5T_NATIVE_DOUBLE, ini->vz, n);
            std::vector<double> birth(n, 0.0), death(n, -1.0);
            for (size_t i = 0; i < n; ++i) {
                if (ini->birth_time_s) birth[i] = ini->birth_time_s[i];
                if (ini->death_time_s) death[i] = ini->death_time_s[i];
            }
            write_1d(h->ions, "birth_time_s", H5T_NATIVE_DOUBLE, birth.data(), n);
            write_1d(h->ions, "death_time_s", H5T_NATIVE_DOUBLE, death.data(), n);
            write_1d(h->ions, "charge_C", H5T_NATIVE_DOUBLE, ini->charge_C, n);
            h->traj = check_id(H5Gcreate2(h->file, "/trajectory", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                               "/trajectory");
            check_rc(H5Fflush(h->file, H5F_SCOPE_GLOBAL), "H5Fflush");
            h->t.reserve(h->batch);
            h->pos.reserve(h->batch * n * 3);
            h->vel.reserve(h->batch * n * 3);
            h->dom.reserve(h->batch * n);
            h->spec.reserve(h->batch * n);
        } catch (...) {
            close_all(h);
            delete h;
            throw;
        }
        *out = h;
    });
}

int icarion_h5_append(icarion_h5* h, double t_s, const double* positions, const double* velocities,
                      const int32_t* domain_index, const uint32_t* species, const double* time_per_ion) {
    return guarded([&] {
        if (!h || h->file < 0) fail(ICARION_ERR_STATE, "writer is closed");
        if (h->n && (!positions || !velocities || !species))
            fail(ICARION_ERR_INVALID, "positions, velocities and species are required");
        const int tp = time_per_ion ? 1 : 0;
        if (h->per_ion_time < 0) h->per_ion_time = tp;
        if (h->per_ion_time != tp)
            fail(ICARION_ERR_INVALID, "time_per_ion must be given for every snapshot or for none");
        if (h->rows == h->batch) flush_rows(h);
        if (h->rows == 0) {
            h->t.clear(); h->pos.clear(); h->vel.clear(); h->dom.clear(); h->spec.clear(); h->tpi.clear();
        }
        const size_t n = h->n;
        h->t.push_back(t_s);
        h->pos.insert(h->pos.end(), positions, positions + 3 * n);
        h->vel.insert(h->vel.end(), velocities, velocities + 3 * n);
        if (domain_index)
            h->dom.insert(h->dom.end(), domain_index, domain_index + n);
        else
            h->dom.insert(h->dom.end(), n, 0);
        h->spec.insert(h->spec.end(), species, species + n);
        if (tp) h->tpi.insert(h->tpi.end(), time_per_ion, time_per_ion + n);
        ++h->rows;
    });
}

int icarion_h5_on_snapshot(void* user, uint64_t, uint64_t, double t_s, size_t n, const double* positions,
                           const double* velocities, const uint32_t* species, const int32_t* domain_index,
                           const uint8_t*) {
    icarion_h5* h = static_cast<icarion_h5*>(user);
    if (!h || n != h->n) {
        g_error = "snapshot ion count differs from the file's";
        return 1;
    }
    return icarion_h5_append(h, t_s, positions, velocities, domain_index, species, nullptr) == ICARION_OK ? 0 : 1;
}

int icarion_h5_flush(icarion_h5* h) {
    return guarded([&] {
        if (!h || h->file < 0) fail(ICARION_ERR_STATE, "writer is closed");
        flush_rows(h);
    });
}

int icarion_h5_update_death_times(icarion_h5* h, const double* death_time_s, size_t n) {
    return guarded([&] {
        if (!h || h->file < 0) fail(ICARION_ERR_STATE, "writer is closed");
        if (n != h->n || (n && !death_time_s)) fail(ICARION_ERR_INVALID, "death times: one per ion");
        Id ds(check_id(H5Dopen2(h->ions, "death_time_s", H5P_DEFAULT), "death_time_s"), H5Dclose);
        if (n) check_rc(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, death_time_s), "death_time_s");
        check_rc(H5Fflush(h->file, H5F_SCOPE_GLOBAL), "H5Fflush");
    });
}

uint64_t icarion_h5_rows(const icarion_h5* h) { return h ? h->written + h->rows : 0; }

int icarion_h5_close(icarion_h5* h) {
    if (!h) return ICARION_OK;
    const int rc = guarded([&] {
        if (h->file >= 0) flush_rows(h);
    });
    close_all(h);
    delete h;
    return rc;
}

}  // extern "C"
