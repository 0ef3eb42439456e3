// wrach_api.hpp -- C++ host mirror of wrach's `runners/api` surface over the
// C ABI in wrach.h (header only; link against libwrach_hip.so).
//
// Names, argument meaning and behaviour follow the reference's Rust API:
//   WrachConfig  runners/bevy/src/config_app.rs:10-35 (+ gravity/dt extension)
//   Particle     runners/bevy/src/state.rs:41-51
//   PackedData   runners/bevy/src/spatial_bin.rs:22-33
//   WrachState   runners/bevy/src/state.rs:17-33 (the parts a caller reads)
//   WrachAPI     runners/api/src/lib.rs:17-88: new, tick, read_data,
//                add_particles, get_simulation_state, and the public
//                `positions` / `velocities` fields
// Where the reference panics, these methods throw wrach::Error.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "wrach.h"

namespace wrach {

struct Vec2 {
    float x = 0.0f;
    float y = 0.0f;
};

inline bool operator==(const Vec2 &a, const Vec2 &b) { return a.x == b.x && a.y == b.y; }

struct Particle {
    Vec2 position;
    Vec2 velocity;
};

struct WrachConfig {
    std::pair<uint16_t, uint16_t> dimensions{480, 352};
    bool boundaries_as_dimensions = false;
    uint16_t cell_size = 3;
    Vec2 gravity{0.0f, 0.0f};  // extension: (0, 0) is the reference's step
    float dt = 1.0f;           // extension: 1 is the reference's step
    int device = 0;
    bool host_pack = false;    // pack uploads on the host (WRACH_FLAG_HOST_PACK)
};

struct PackedData {
    std::vector<uint32_t> indices;
    std::vector<Vec2> positions;   // capacity long, zero padded
    std::vector<Vec2> velocities;  // capacity long, zero padded
};

struct WrachState {
    WrachConfig config;
    wrach_world_settings shader_settings{};
    PackedData packed_data;
};

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc != WRACH_OK) throw Error(rc, std::string("wrach: ") + wrach_last_error());
}

class WrachAPI {
  public:
    // All the positions / velocities of the particles, as (x, y) pairs.
    std::vector<std::pair<float, float>> positions;
    std::vector<std::pair<float, float>> velocities;

    explicit WrachAPI(WrachConfig config = WrachConfig()) {
        wrach_config c;
        wrach_config_default(&c);
        c.width = config.dimensions.first;
        c.height = config.dimensions.second;
        c.cell_size = config.cell_size;
        c.boundaries_as_dimensions = config.boundaries_as_dimensions ? 1 : 0;
        c.gravity[0] = config.gravity.x;
        c.gravity[1] = config.gravity.y;
        c.dt = config.dt;
        c.device = config.device;
        c.flags = config.host_pack ? WRACH_FLAG_HOST_PACK : 0u;
        check(wrach_create(&c, &ctx_));
        state_.config = config;
        check(wrach_settings(ctx_, &state_.shader_settings));
    }

    ~WrachAPI() { wrach_destroy(ctx_); }
    WrachAPI(const WrachAPI &) = delete;
    WrachAPI &operator=(const WrachAPI &) = delete;

    // Run a single tick/frame of the simulation, then read_data().
    void tick() {
        check(wrach_step(ctx_, 1));
        readback();
        read_data();
    }

    // Copy the last read-back frame into `positions` / `velocities`.
    void read_data() {
        positions.clear();
        velocities.clear();
        positions.reserve(state_.packed_data.positions.size());
        velocities.reserve(state_.packed_data.velocities.size());
        for (const Vec2 &p : state_.packed_data.positions) positions.emplace_back(p.x, p.y);
        for (const Vec2 &v : state_.packed_data.velocities) velocities.emplace_back(v.x, v.y);
    }

    // Add particles to the simulation (re-packs the whole store and uploads).
    void add_particles(const std::vector<Particle> &particles) {
        std::vector<float> pos, vel;
        pos.reserve(2 * particles.size());
        vel.reserve(2 * particles.size());
        for (const Particle &p : particles) {
            pos.push_back(p.position.x);
            pos.push_back(p.position.y);
            vel.push_back(p.velocity.x);
            vel.push_back(p.velocity.y);
        }
        check(wrach_add_particles(ctx_, pos.data(), vel.data(), (uint32_t)particles.size()));
        check(wrach_settings(ctx_, &state_.shader_settings));
    }

    const WrachState &get_simulation_state() const { return state_; }

    // Not in the reference: run n frames without a readback.
    void step(uint32_t nframes) { check(wrach_step(ctx_, nframes)); }

    // ParticleStore::update_from_gpu (a TODO in particle_store.rs:74-85): the
    // store becomes the simulated particles, so add_particles continues from
    // them.
    void update_store_from_gpu() { check(wrach_store_write_back(ctx_)); }

    // Restores a packed triple saved from get_simulation_state() after a tick
    // (only its first indices.back() particles are read).
    void load_packed(const PackedData &d) {
        const uint32_t n = d.indices.empty() ? 0u : d.indices.back();
        if (d.positions.size() < n || d.velocities.size() < n)
            throw Error(WRACH_E_BUFFER, "wrach: packed arrays shorter than the last offset");
        check(wrach_load_packed(ctx_, d.indices.data(), d.indices.size(),
                                reinterpret_cast<const float *>(d.positions.data()),
                                reinterpret_cast<const float *>(d.velocities.data()), n));
        check(wrach_settings(ctx_, &state_.shader_settings));
    }
    // The active window over the store (particle_store.rs:19-26): move the
    // view; the reference keeps it at the origin (builder.rs:61).
    void set_view_anchor(float x, float y) {
        check(wrach_set_view_anchor(ctx_, x, y));
        check(wrach_settings(ctx_, &state_.shader_settings));
    }
    // Not in the reference: the solver (WRACH_SOLVER_R, or WRACH_SOLVER_S with
    // separation passes over the 3x3 cells before the reference step) and
    // the last mode S frame's overlap before and after its passes.
    void set_solver(uint32_t mode, uint32_t max_passes = 0, float overlap_tol = 0.0f) {
        check(wrach_set_solver(ctx_, mode, max_passes, overlap_tol));
    }
    wrach_solver_report solver_stats() {
        wrach_solver_report r{};
        check(wrach_solver_stats(ctx_, &r));
        return r;
    }
    wrach_ctx *raw() { return ctx_; }

  private:
    void readback() {
        uint32_t cap = 0, nidx = 0, n = 0;
        check(wrach_capacity(ctx_, &cap, &nidx));
        PackedData &d = state_.packed_data;
        d.indices.assign(nidx, 0);
        d.positions.assign(cap, Vec2{});
        d.velocities.assign(cap, Vec2{});
        check(wrach_read(ctx_, reinterpret_cast<float *>(d.positions.data()),
                         reinterpret_cast<float *>(d.velocities.data()), cap, d.indices.data(),
                         nidx, &n));
        check(wrach_settings(ctx_, &state_.shader_settings));
    }

    wrach_ctx *ctx_ = nullptr;
    WrachState state_;
};

static_assert(sizeof(Vec2) == 2 * sizeof(float), "Vec2 is read as interleaved floats");

}  // namespace wrach
