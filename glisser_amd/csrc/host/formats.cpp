// Configuration, state and track files (layouts of src/data.cpp:146-877).
#include "formats.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <functional>
#include <iomanip>
#include <istream>
#include <limits>
#include <map>
#include <ostream>
#include <sstream>
#include <stdexcept>

#include "convert.h"
#include "vec3.h"

namespace glx {
namespace host {

namespace {

uint32_t parse_u32(const std::string& s) {
    size_t used = 0;
    const unsigned long v = std::stoul(s, &used);
    if (v > 0xFFFFFFFFul) throw std::out_of_range("u32");
    return (uint32_t)v;
}

bool parse_flag(const std::string& s) { return std::stoi(s) != 0; }

using Setter = std::function<void(Configuration&, const std::string&)>;

// Key -> how its value is stored (the keys are the reference's file format).
const std::map<std::string, Setter>& config_keys() {
    static const std::map<std::string, Setter> keys = {
        {"Initial-Time", [](Configuration& c, const std::string& v) { c.t_0 = std::stod(v); }},
        {"Time-Step", [](Configuration& c, const std::string& v) { c.dt = std::stod(v); }},
        {"Final-Time", [](Configuration& c, const std::string& v) { c.t_f = std::stod(v); }},
        {"Time-Block-Size", [](Configuration& c, const std::string& v) { c.tbsize = parse_u32(v); }},
        {"Big-G", [](Configuration& c, const std::string& v) { c.big_g = std::stod(v); }},
        {"Encounter-RH-Factor", [](Configuration& c, const std::string& v) { c.encounter_sphere_factor = std::stod(v); }},
        {"Cull-Radius", [](Configuration& c, const std::string& v) { c.cull_radius = std::stod(v); }},
        {"Outer-Limit", [](Configuration& c, const std::string& v) { c.outer_radius = std::stod(v); }},
        {"Limit-Particle-Count", [](Configuration& c, const std::string& v) { c.max_particle = parse_u32(v); }},
        {"Log-Interval", [](Configuration& c, const std::string& v) { c.print_every = parse_u32(v); }},
        {"Resync-Interval", [](Configuration& c, const std::string& v) { c.resync_every = parse_u32(v); }},
        {"Read-Planet-History", [](Configuration& c, const std::string& v) { c.interp_planets = parse_flag(v); }},
        {"Planet-History-Max-Planet-Count", [](Configuration& c, const std::string& v) { c.interp_maxpl = parse_u32(v); }},
        {"Planet-History-File", [](Configuration& c, const std::string& v) { c.planet_history_file = v; }},
        {"Swift-Path", [](Configuration& c, const std::string& v) { c.swift_path = v; }},
        {"Track-Interval", [](Configuration& c, const std::string& v) { c.track_every = parse_u32(v); }},
        {"Swift-History-Interval", [](Configuration& c, const std::string& v) { c.swift_hist_every = parse_u32(v); }},
        {"Swift-Process-Count", [](Configuration& c, const std::string& v) { c.num_swift = parse_u32(v); }},
        {"Swift-Process-Min-Particle-Count", [](Configuration& c, const std::string& v) { c.swift_part_min = parse_u32(v); }},
        {"Swift-Status-Length", [](Configuration& c, const std::string& v) { c.swift_statlen = parse_u32(v); }},
        {"Write-Barycentric-Track", [](Configuration& c, const std::string& v) { c.write_bary_track = parse_flag(v); }},
        {"Split-Track-File", [](Configuration& c, const std::string& v) { c.split_track_file = parse_u32(v); }},
        {"Dump-Interval", [](Configuration& c, const std::string& v) { c.dump_every = parse_u32(v); }},
        {"Resolve-Encounters", [](Configuration& c, const std::string& v) { c.resolve_encounters = parse_flag(v); }},
        {"Write-Split-Output", [](Configuration& c, const std::string& v) { c.writesplit = parse_flag(v); }},
        {"Read-Split-Input", [](Configuration& c, const std::string& v) { c.readsplit = parse_flag(v); }},
        {"Write-Binary-Output", [](Configuration& c, const std::string& v) { c.writebinary = parse_flag(v); }},
        {"Read-Binary-Input", [](Configuration& c, const std::string& v) { c.readbinary = parse_flag(v); }},
        {"Input-File", [](Configuration& c, const std::string& v) { c.hybridin = v; }},
        {"Output-File", [](Configuration& c, const std::string& v) { c.hybridout = v; }},
        {"Particle-Input-File", [](Configuration& c, const std::string& v) { c.icsin = v; }},
        {"Planet-Input-File", [](Configuration& c, const std::string& v) { c.plin = v; }},
        {"Output-Folder", [](Configuration& c, const std::string& v) { c.outfolder = v; }},
        {"Read-Input-Momenta", [](Configuration& c, const std::string& v) { c.readmomenta = parse_flag(v); }},
        {"Write-Output-Momenta", [](Configuration& c, const std::string& v) { c.writemomenta = parse_flag(v); }},
    };
    return keys;
}

template <typename T>
void put(std::ostream& out, T v) {
    out.write(reinterpret_cast<const char*>(&v), sizeof v);
}

template <typename T>
T get(std::istream& in) {
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof v);
    if (!in) throw std::runtime_error("state file ends early");
    return v;
}

void put_vec_text(std::ostream& out, const double* p) { out << p[0] << " " << p[1] << " " << p[2] << "\n"; }

std::istringstream next_line(std::istream& in) {
    std::string s;
    if (!std::getline(in, s)) throw std::runtime_error("state file ends early");
    return std::istringstream(s);
}

void resize_particles(State& s, size_t n) {
    s.r.assign(3 * n, 0.0);
    s.v.assign(3 * n, 0.0);
    s.id.assign(n, 0);
    s.flags.assign(n, 0);
    s.deathtime.assign(n, 0.0);
}

void resize_planets(State& s, size_t n) {
    s.pm.assign(n, 0.0);
    s.pr.assign(3 * n, 0.0);
    s.pv.assign(3 * n, 0.0);
    s.pid.assign(n, 0);
}

// Planet velocity as written: times its GM when writing momenta.
double vel_scale(const State& s, size_t i, bool momenta) { return momenta ? s.pm[i] : 1.0; }

void put_track_header_and_planets(std::ostream& out, double time, const State& s, bool elements, bool bary) {
    put<double>(out, time);
    const size_t np = s.n_planets();
    put<uint64_t>(out, np > 0 ? (uint64_t)(np - 1) : 0);
    for (size_t i = 1; i < np; i++) {
        float x[6];
        if (elements) {
            // the Sun, or the barycentre of every body but planet i
            double cm = s.pm[0];
            Vec3 cr = Vec3::load(s.pr.data(), 0) * s.pm[0];
            Vec3 cv = Vec3::load(s.pv.data(), 0) * s.pm[0];
            if (bary)
                for (size_t j = 1; j < np; j++)
                    if (j != i) {
                        cm += s.pm[j];
                        cr += Vec3::load(s.pr.data(), j) * s.pm[j];
                        cv += Vec3::load(s.pv.data(), j) * s.pm[j];
                    }
            cr = cr / cm;
            cv = cv / cm;
            const Elements e = to_elements(s.pm[i] + cm, Vec3::load(s.pr.data(), i) - cr,
                                           Vec3::load(s.pv.data(), i) - cv);
            const double el[6] = {e.a, e.e, e.i, e.capom, e.om, e.f};
            for (int k = 0; k < 6; k++) x[k] = (float)el[k];
        } else {
            for (int k = 0; k < 3; k++) {
                x[k] = (float)s.pr[3 * i + k];
                x[3 + k] = (float)s.pv[3 * i + k];
            }
        }
        put<uint32_t>(out, s.pid[i]);
        for (float f : x) put<float>(out, f);
    }
}

}  // namespace

void read_configuration(std::istream& in, Configuration* out) {
    const auto& keys = config_keys();
    std::string line;
    size_t num = 0;
    while (std::getline(in, line)) {
        num++;
        if (line.empty() || line[0] == '#') continue;
        const size_t sp = line.find(' ');
        if (sp == std::string::npos)
            throw std::runtime_error("Unrecognized line " + std::to_string(num) + ": " + line);
        const std::string key = line.substr(0, sp), value = line.substr(sp + 1);
        auto it = keys.find(key);
        try {
            if (it == keys.end()) throw std::invalid_argument(key);
            it->second(*out, value);
        } catch (const std::logic_error&) {  // invalid_argument, out_of_range
            throw std::runtime_error("Could not parse line " + std::to_string(num) + ": " + line);
        }
    }
    if (out->outfolder.empty()) out->outfolder = "./";
    if (!out->writesplit && out->hybridout.empty()) throw std::runtime_error("Error: Output-File was not specified");
    if (out->resync_every != 1 && out->resolve_encounters)
        throw std::runtime_error("Error: Resync-Interval must be 1 when Resolve-Encounters is enabled");
    if (!out->writesplit && out->hybridin.empty()) throw std::runtime_error("Error: Input-File was not specified");
    if (out->readsplit && (out->plin.empty() || out->icsin.empty()))
        throw std::runtime_error(
            "Error: Read-Split-Input was selected but Particle-Input-File or Planet-Input-File were not specified");
    if (!out->interp_planets && out->resolve_encounters)
        throw std::runtime_error("Error: Cannot resolve encounters if not using an interpolated planetary history file");
}

void write_configuration(std::ostream& o, const Configuration& c) {
    o << std::setprecision(17);
    o << "Initial-Time " << c.t_0 << "\n"
      << "Time-Step " << c.dt << "\n"
      << "Final-Time " << c.t_f << "\n"
      << "Time-Block-Size " << c.tbsize << "\n"
      << "Encounter-RH-Factor " << c.encounter_sphere_factor << "\n"
      << "Cull-Radius " << c.cull_radius << "\n"
      << "Outer-Limit " << c.outer_radius << "\n"
      << "Limit-Particle-Count " << c.max_particle << "\n"
      << "Log-Interval " << c.print_every << "\n"
      << "Track-Interval " << c.track_every << "\n"
      << "Resync-Interval " << c.resync_every << "\n"
      << "Swift-Path " << c.swift_path << "\n"
      << "Write-Barycentric-Track " << c.write_bary_track << "\n"
      << "Split-Track-File " << c.split_track_file << "\n"
      << "Dump-Interval " << c.dump_every << "\n"
      << "Resolve-Encounters " << c.resolve_encounters << "\n"
      << "Write-Split-Output " << c.writesplit << "\n"
      << "Write-Binary-Output " << c.writebinary << "\n"
      << "Read-Split-Input " << c.readsplit << "\n"
      << "Read-Binary-Input " << c.readbinary << "\n"
      << "Input-File " << c.hybridin << "\n"
      << "Output-File " << c.hybridout << "\n"
      << "Particle-Input-File " << c.icsin << "\n"
      << "Planet-Input-File " << c.plin << "\n"
      << "Output-Folder " << c.outfolder << "\n"
      << "Read-Input-Momenta " << c.readmomenta << "\n"
      << "Write-Output-Momenta " << c.writemomenta << "\n"
      << "Read-Planet-History " << c.interp_planets << "\n"
      << "Planet-History-Max-Planet-Count " << c.interp_maxpl << "\n"
      << "Swift-History-Interval " << c.swift_hist_every << "\n"
      << "Swift-Process-Count " << c.num_swift << "\n"
      << "Swift-Process-Min-Particle-Count " << c.swift_part_min << "\n"
      << "Swift-Status-Length " << c.swift_statlen << "\n"
      << "Planet-History-File " << c.planet_history_file << "\n"
      << "Big-G " << c.big_g << "\n";
    o.flush();
}

void save_state_ascii(std::ostream& out, const State& s, bool momenta) {
    out << s.n_planets() << "\n" << std::setprecision(17);
    for (size_t i = 0; i < s.n_planets(); i++) {
        const double k = vel_scale(s, i, momenta);
        const double v[3] = {s.pv[3 * i] * k, s.pv[3 * i + 1] * k, s.pv[3 * i + 2] * k};
        out << s.pm[i] << "\n";
        put_vec_text(out, &s.pr[3 * i]);
        put_vec_text(out, v);
        out << s.pid[i] << "\n";
    }
    out << s.n_particles() << "\n";
    for (size_t i = 0; i < s.n_particles(); i++) {
        put_vec_text(out, &s.r[3 * i]);
        put_vec_text(out, &s.v[3 * i]);
        out << s.id[i] << " " << s.flags[i] << " " << s.deathtime[i] << "\n";
    }
    out.flush();
}

void save_state_binary(std::ostream& out, const State& s, bool momenta) {
    put<uint64_t>(out, s.n_planets());
    for (size_t i = 0; i < s.n_planets(); i++) {
        const double k = vel_scale(s, i, momenta);
        put<uint32_t>(out, s.pid[i]);
        put<double>(out, s.pm[i]);
        for (int c = 0; c < 3; c++) put<double>(out, s.pr[3 * i + c]);
        for (int c = 0; c < 3; c++) put<double>(out, s.pv[3 * i + c] * k);
    }
    put<uint64_t>(out, s.n_particles());
    for (size_t i = 0; i < s.n_particles(); i++) {
        put<uint32_t>(out, s.id[i]);
        for (int c = 0; c < 3; c++) put<double>(out, s.r[3 * i + c]);
        for (int c = 0; c < 3; c++) put<double>(out, s.v[3 * i + c]);
        put<uint16_t>(out, s.flags[i]);
        put<double>(out, s.deathtime[i]);
    }
    out.flush();
}

void save_state_split(std::ostream& pl, std::ostream& pa, const State& s, bool momenta) {
    pl << s.n_planets() << "\n" << std::setprecision(17);
    for (size_t i = 0; i < s.n_planets(); i++) {
        const double k = vel_scale(s, i, momenta);
        const double v[3] = {s.pv[3 * i] * k, s.pv[3 * i + 1] * k, s.pv[3 * i + 2] * k};
        pl << s.pm[i] << "\n";
        put_vec_text(pl, &s.pr[3 * i]);
        put_vec_text(pl, v);
    }
    pa << s.n_particles() << "\n" << std::setprecision(17);
    for (size_t i = 0; i < s.n_particles(); i++) {
        put_vec_text(pa, &s.r[3 * i]);
        put_vec_text(pa, &s.v[3 * i]);
        pa << s.deathtime[i] << " " << s.flags[i] << " " << s.id[i] << "\n";
    }
    pl.flush();
    pa.flush();
}

State load_state_ascii(std::istream& in, uint64_t max_particles) {
    State s;
    size_t np = 0;
    next_line(in) >> np;
    resize_planets(s, np);
    for (size_t i = 0; i < np; i++) {
        next_line(in) >> s.pm[i];
        next_line(in) >> s.pr[3 * i] >> s.pr[3 * i + 1] >> s.pr[3 * i + 2];
        next_line(in) >> s.pv[3 * i] >> s.pv[3 * i + 1] >> s.pv[3 * i + 2];
        next_line(in) >> s.pid[i];
    }
    size_t n = 0;
    next_line(in) >> n;
    n = (size_t)std::min<uint64_t>(n, max_particles);
    resize_particles(s, n);
    for (size_t i = 0; i < n; i++) {
        next_line(in) >> s.r[3 * i] >> s.r[3 * i + 1] >> s.r[3 * i + 2];
        next_line(in) >> s.v[3 * i] >> s.v[3 * i + 1] >> s.v[3 * i + 2];
        auto ls = next_line(in);
        ls >> s.id[i] >> s.flags[i] >> s.deathtime[i];
    }
    return s;
}

State load_state_binary(std::istream& in, uint64_t max_particles) {
    State s;
    const size_t np = (size_t)get<uint64_t>(in);
    resize_planets(s, np);
    for (size_t i = 0; i < np; i++) {
        s.pid[i] = get<uint32_t>(in);
        s.pm[i] = get<double>(in);
        for (int c = 0; c < 3; c++) s.pr[3 * i + c] = get<double>(in);
        for (int c = 0; c < 3; c++) s.pv[3 * i + c] = get<double>(in);
    }
    const size_t n = (size_t)std::min<uint64_t>(get<uint64_t>(in), max_particles);
    resize_particles(s, n);
    for (size_t i = 0; i < n; i++) {
        s.id[i] = get<uint32_t>(in);
        for (int c = 0; c < 3; c++) s.r[3 * i + c] = get<double>(in);
        for (int c = 0; c < 3; c++) s.v[3 * i + c] = get<double>(in);
        s.flags[i] = get<uint16_t>(in);
        s.deathtime[i] = get<double>(in);
    }
    return s;
}

State load_state_split(std::istream& pl, std::istream& pa, uint64_t max_particles) {
    State s;
    size_t np = 0;
    pl >> np;
    resize_planets(s, np);
    for (size_t i = 0; i < np; i++) {
        pl >> s.pm[i];
        pl >> s.pr[3 * i] >> s.pr[3 * i + 1] >> s.pr[3 * i + 2];
        pl >> s.pv[3 * i] >> s.pv[3 * i + 1] >> s.pv[3 * i + 2];
        s.pid[i] = (uint32_t)i;
    }
    size_t n = 0;
    pa >> n;
    n = (size_t)std::min<uint64_t>(n, max_particles);
    resize_particles(s, n);
    for (size_t i = 0; i < n; i++) {
        pa >> s.r[3 * i] >> s.r[3 * i + 1] >> s.r[3 * i + 2];
        pa >> s.v[3 * i] >> s.v[3 * i + 1] >> s.v[3 * i + 2];
        // optional "deathtime flags id" tail; anything else on the line means
        // a fresh particle with id = its index
        std::string tok;
        pa >> tok;
        if (tok.empty() || !std::isdigit((unsigned char)tok[0])) {
            pa.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
            s.deathtime[i] = 0;
            s.id[i] = (uint32_t)i;
            s.flags[i] = 0;
        } else {
            s.deathtime[i] = (double)std::stof(tok);
            pa >> s.flags[i] >> s.id[i];
        }
    }
    if (!pa && !pa.eof()) throw std::runtime_error("particle input file is malformed");
    return s;
}

void apply_input_units(State& s, bool read_momenta, double big_g) {
    for (size_t i = 0; i < s.n_planets(); i++) {
        if (read_momenta)
            for (int c = 0; c < 3; c++) s.pv[3 * i + c] /= s.pm[i];
        s.pm[i] *= big_g;
    }
}

void track_center(const State& s, bool bary, double* mass, double r[3], double v[3]) {
    double cm = s.n_planets() ? s.pm[0] : 0.0;
    Vec3 cr = s.n_planets() ? Vec3::load(s.pr.data(), 0) : Vec3();
    Vec3 cv = s.n_planets() ? Vec3::load(s.pv.data(), 0) : Vec3();
    if (bary) {
        Vec3 rs, vs;
        double ms = 0;
        for (size_t j = 0; j < s.n_planets(); j++) {
            rs += Vec3::load(s.pr.data(), j) * s.pm[j];
            vs += Vec3::load(s.pv.data(), j) * s.pm[j];
            ms += s.pm[j];
        }
        cr = rs / ms;
        cv = vs / ms;
        cm = ms;
    }
    *mass = cm;
    cr.store(r, 0);
    cv.store(v, 0);
}

void write_track_snapshot(std::ostream& out, double time, const State& s, bool to_elements, bool bary) {
    put_track_header_and_planets(out, time, s, to_elements, bary);
    const size_t n = s.n_particles();
    put<uint64_t>(out, n);
    double cm = 0, crv[3], cvv[3];
    track_center(s, to_elements && bary, &cm, crv, cvv);
    const Vec3 cr(crv[0], crv[1], crv[2]), cv(cvv[0], cvv[1], cvv[2]);
    for (size_t i = 0; i < n; i++) {
        TrackRecord rec;
        rec.id = s.id[i];
        if (to_elements) {
            const Elements e = host::to_elements(cm, Vec3::load(s.r.data(), i) - cr, Vec3::load(s.v.data(), i) - cv);
            const double el[6] = {e.a, e.e, e.i, e.capom, e.om, e.f};
            for (int k = 0; k < 6; k++) rec.x[k] = (float)el[k];
        } else {
            for (int k = 0; k < 3; k++) {
                rec.x[k] = (float)s.r[3 * i + k];
                rec.x[3 + k] = (float)s.v[3 * i + k];
            }
        }
        out.write(reinterpret_cast<const char*>(&rec), sizeof rec);
    }
    out.flush();
}

void write_track_records(std::ostream& out, double time, const State& planets, bool to_elements, bool bary,
                         const TrackRecord* records, uint64_t n) {
    put_track_header_and_planets(out, time, planets, to_elements, bary);
    put<uint64_t>(out, n);
    out.write(reinterpret_cast<const char*>(records), (std::streamsize)(n * sizeof(TrackRecord)));
    out.flush();
}

}  // namespace host
}  // namespace glx
