// The reference's API-level tests, written against the C++ host mirror
// (include/wrach_api.hpp).  Built and run by tests/test_cpp_api.py.
#include <cstdio>
#include <vector>

#include "wrach_api.hpp"

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

// runners/api/src/lib.rs:101-126 test_api_returns_data
static void api_returns_data() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {10, 10};
    cfg.cell_size = 3;
    wrach::WrachAPI wrach(cfg);
    std::vector<wrach::Particle> particles(3, wrach::Particle{{5.0f, 5.0f}, {0.5f, 0.5f}});
    wrach.add_particles(particles);
    for (int i = 0; i < 5; ++i) wrach.tick();
    CHECK(wrach.positions.size() == 164);
    CHECK(wrach.positions[0] != std::make_pair(0.0f, 0.0f));
    CHECK(wrach.velocities.size() == 164);
    CHECK(wrach.velocities[0] != std::make_pair(0.0f, 0.0f));
}

// runners/bevy/src/compute/03_prefix_sum.rs:150-207 prefix_sum_for_small_arrays
static void prefix_sum_for_small_arrays() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {10, 10};
    cfg.cell_size = 5;
    wrach::WrachAPI wrach(cfg);
    wrach.add_particles({{{0.1f, 0.1f}, {0, 0}}, {{2.0f, 2.0f}, {0, 0}},
                         {{5.0f, 5.0f}, {0, 0}}, {{10.0f, 10.0f}, {0, 0}}});
    for (int i = 0; i < 4; ++i) wrach.tick();
    const std::vector<uint32_t> want = {0, 0, 2, 2, 2, 2, 3, 3, 3, 3, 4};
    CHECK(wrach.get_simulation_state().packed_data.indices == want);
}

// runners/bevy/src/compute/04_pack_particle_data.rs:72-139 packed_data
static void packed_data() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {10, 10};
    cfg.cell_size = 3;
    wrach::WrachAPI wrach(cfg);
    wrach.add_particles({{{0.1f, 0.1f}, {0, 0}}, {{2.5f, 2.5f}, {0, 0}},
                         {{5.0f, 5.0f}, {0, 0}}, {{10.0f, 10.0f}, {0, 0}}});
    for (int i = 0; i < 4; ++i) wrach.tick();
    const auto &pos = wrach.get_simulation_state().packed_data.positions;
    CHECK((pos[0] == wrach::Vec2{0.1f, 0.1f}));
    CHECK((pos[1] == wrach::Vec2{2.5f, 2.5f}));
    CHECK((pos[2] == wrach::Vec2{5.0f, 5.0f}));
    CHECK((pos[3] == wrach::Vec2{10.0f, 10.0f}));
}

// Where the reference would panic on a full buffer, the mirror throws.
static void capacity_error() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {6, 6};
    cfg.cell_size = 2;
    wrach::WrachAPI wrach(cfg);
    bool threw = false;
    try {
        wrach.add_particles(std::vector<wrach::Particle>(75, wrach::Particle{{1, 1}, {0, 0}}));
    } catch (const wrach::Error &e) {
        threw = e.code() == WRACH_E_CAPACITY;
    }
    CHECK(threw);
    CHECK(wrach.get_simulation_state().shader_settings.particles_in_frame_count == 0);
}

// Not in the reference: a checkpoint of the packed triple restored into a
// second world continues like the first; the written-back store re-packs to
// the simulated state (host store and device store).
static void checkpoint_and_write_back(bool host_pack) {
    wrach::WrachConfig cfg;
    cfg.dimensions = {30, 20};
    cfg.cell_size = 1;
    cfg.gravity = {0.0f, -0.1f};
    cfg.host_pack = host_pack;
    std::vector<wrach::Particle> ps;
    for (int i = 0; i < 300; ++i)
        ps.push_back({{(float)(i % 29) + 0.3f, (float)(i / 29) * 1.7f + 0.2f},
                      {(float)((i * 7) % 5) * 0.2f - 0.4f, (float)((i * 3) % 7) * 0.1f - 0.3f}});
    wrach::WrachAPI a(cfg), b(cfg);
    a.add_particles(ps);
    for (int i = 0; i < 6; ++i) a.tick();
    const wrach::PackedData saved = a.get_simulation_state().packed_data;
    b.load_packed(saved);
    for (int i = 0; i < 4; ++i) {
        a.tick();
        b.tick();
    }
    CHECK(a.get_simulation_state().packed_data.indices == b.get_simulation_state().packed_data.indices);
    CHECK(a.positions == b.positions);
    CHECK(a.velocities == b.velocities);
    a.update_store_from_gpu();
    a.add_particles({});  // re-pack of the written-back store
    a.step(0);
    b.step(0);
    a.tick();
    b.tick();
    CHECK(a.positions == b.positions);
    CHECK(a.velocities == b.velocities);
}

// Mode S through the mirror: two particles 0.6 apart in neighbouring cells
// (which mode R never separates) end a pass exactly 1 apart before the
// step moves them; with no velocity and no gravity the frame keeps that.
static void stencil_solver() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {16, 16};
    cfg.cell_size = 1;
    wrach::WrachAPI a(cfg);
    a.add_particles({{{5.5f, 5.5f}, {0.0f, 0.0f}}, {{6.1f, 5.5f}, {0.0f, 0.0f}}});
    a.set_solver(WRACH_SOLVER_S, 2, 0.0f);
    a.tick();
    CHECK(a.positions[1].first - a.positions[0].first == 1.0f);
    const auto st = a.solver_stats();
    CHECK(st.passes == 1);  // the second pass starts with no overlap left
    CHECK(st.max_overlap_after == 0.0f && st.max_overlap_before > 0.39f && st.max_overlap_before < 0.41f);
    CHECK(st.pairs_before == 2 && st.pairs_after == 2);  // the pair, from both sides, now touching
    CHECK(st.overlap_sum_before == 2 * 6710888u);  // 1 - d = 0.4000001: 6,710,888 units of 2^-24, twice
    CHECK(st.overlap_sum_after == 0);
    a.set_solver(WRACH_SOLVER_R);
    a.tick();
    CHECK(a.positions[1].first - a.positions[0].first == 1.0f);
}

// The movable view through the mirror: two particles in the first view, three
// beyond it; moving the view packs the three (and keeps the two in the
// store), moving back packs the two again.
static void movable_view() {
    wrach::WrachConfig cfg;
    cfg.dimensions = {16, 16};
    cfg.cell_size = 1;
    wrach::WrachAPI a(cfg);
    a.add_particles({{{3.5f, 3.5f}, {0.0f, 0.0f}}, {{9.5f, 2.5f}, {0.0f, 0.0f}},
                     {{20.5f, 4.5f}, {0.0f, 0.0f}}, {{25.5f, 9.5f}, {0.0f, 0.0f}},
                     {{30.5f, 1.5f}, {0.0f, 0.0f}}});
    a.tick();
    CHECK(a.get_simulation_state().shader_settings.particles_in_frame_count == 2);
    a.set_view_anchor(16.0f, 0.0f);
    CHECK(a.get_simulation_state().shader_settings.particles_in_frame_count == 3);
    CHECK(a.get_simulation_state().shader_settings.view_anchor[0] == 16.0f);
    a.tick();
    // row-major cells of the new view: rows 1, 4 and 9
    CHECK(a.positions[0] == std::make_pair(30.5f, 1.5f) && a.positions[1] == std::make_pair(20.5f, 4.5f) &&
          a.positions[2] == std::make_pair(25.5f, 9.5f));
    a.set_view_anchor(0.0f, 0.0f);
    a.tick();
    CHECK(a.get_simulation_state().shader_settings.particles_in_frame_count == 2);
    CHECK(a.positions[0] == std::make_pair(9.5f, 2.5f) && a.positions[1] == std::make_pair(3.5f, 3.5f));
}

int main() {
    api_returns_data();
    prefix_sum_for_small_arrays();
    packed_data();
    capacity_error();
    checkpoint_and_write_back(false);
    checkpoint_and_write_back(true);
    stencil_solver();
    movable_view();
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
