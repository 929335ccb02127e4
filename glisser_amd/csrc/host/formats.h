// File formats of the host API: the configuration file, the state files
// (ASCII, binary, split planet/particle files) and the binary track, with
// the reference's layouts (src/data.cpp:146-877, README.md) so files move
// between the two programs unchanged.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <string>
#include <vector>

namespace glx {
namespace host {

// Run configuration (src/data.h:368-414); defaults of src/data.cpp:146-186.
struct Configuration {
    double t_0 = 0, t_f = 1e4, dt = 1, big_g = 1;
    uint32_t tbsize = 1024, print_every = 10, dump_every = 1000, track_every = 0;
    uint32_t max_particle = 0xFFFFFFFFu;
    double encounter_sphere_factor = 0;
    uint32_t split_track_file = 500000000;
    double outer_radius = 300;
    uint32_t resync_every = 1, swift_hist_every = 0, num_swift = 1, swift_part_min = 10, swift_statlen = 13;
    bool write_bary_track = true, interp_planets = false;
    uint32_t interp_maxpl = 16;
    std::string planet_history_file;
    double cull_radius = 0.5;
    bool resolve_encounters = false, readmomenta = false, writemomenta = false;
    bool readsplit = false, writesplit = false, writebinary = false, readbinary = false;
    std::string icsin, plin, hybridin, hybridout = "state.out", outfolder = "output/", swift_path;
};

// Parses "Key value" lines ('#' comments, blank lines skipped) and applies the
// reference's consistency rules; throws std::runtime_error naming the line.
void read_configuration(std::istream& in, Configuration* out);
// Every key in the reference's order, doubles with 17 significant digits.
void write_configuration(std::ostream& out, const Configuration& c);

// Planets and particles of a state file.  r, v: n x 3.
struct State {
    std::vector<double> pm, pr, pv;  // planets: GM, position, velocity
    std::vector<uint32_t> pid;
    std::vector<double> r, v, deathtime;  // particles
    std::vector<uint32_t> id;
    std::vector<uint16_t> flags;
    size_t n_planets() const { return pm.size(); }
    size_t n_particles() const { return id.size(); }
};

enum class StateFormat { Ascii, Binary, Split };

// Writers (src/data.cpp:586-669).  write_momenta multiplies planet velocities
// by their GM.
void save_state_ascii(std::ostream& out, const State& s, bool write_momenta);
void save_state_binary(std::ostream& out, const State& s, bool write_momenta);
void save_state_split(std::ostream& planets, std::ostream& particles, const State& s, bool write_momenta);
// Readers (src/data.cpp:374-525), then the momenta and Big-G conversions of
// load_data (src/data.cpp:527-584).  At most max_particles particles are read.
// The binary reader takes the deathtime as the double the writer stores.
State load_state_ascii(std::istream& in, uint64_t max_particles);
State load_state_binary(std::istream& in, uint64_t max_particles);
State load_state_split(std::istream& planets, std::istream& particles, uint64_t max_particles);
void apply_input_units(State& s, bool read_momenta, double big_g);

// One snapshot of the binary track (src/data.cpp:718-815): time, planets
// 1..n-1, then the given particles in the given order, each as a u32 id and
// six float32 values -- orbital elements (a, e, i, Omega, omega, f) about the
// Sun or, with barycentric, about the barycentre of the other bodies, or the
// Cartesian state.  Planet 0 (the Sun) is the centre and is not written.
void write_track_snapshot(std::ostream& out, double time, const State& s, bool to_elements, bool barycentric);

// Centre of the particle elements of a track snapshot: the Sun (GM m[0]),
// or with barycentric the barycentre of all bodies with their total GM.
void track_center(const State& planets, bool barycentric, double* mass, double r[3], double v[3]);

// 28-byte particle record of a track snapshot.
#pragma pack(push, 1)
struct TrackRecord {
    uint32_t id;
    float x[6];
};
#pragma pack(pop)
static_assert(sizeof(TrackRecord) == 28, "track record is 28 bytes");

// A track snapshot whose particle records were formed elsewhere (the device
// packer): time, planets as above, then n records.
void write_track_records(std::ostream& out, double time, const State& planets, bool to_elements, bool barycentric,
                         const TrackRecord* records, uint64_t n);

}  // namespace host
}  // namespace glx
