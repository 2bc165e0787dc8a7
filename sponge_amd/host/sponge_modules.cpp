// Implementation of the C++ module facades (see sponge_modules.hpp).
#include "sponge_modules.hpp"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace sponge_amd {

namespace {

void default_handler(int code, const char* where, const char* reason) {
    fprintf(stderr, "\nSPONGE ERROR %d in %s\n%s\n", code, where, reason ? reason : "");
    exit(code);
}

ErrorHandler g_handler = default_handler;

void check(int rc, const char* where) {
    if (rc == 0) return;
    char reason[128];
    if (rc < 1000)
        snprintf(reason, sizeof(reason), "Reason:\n\tHIP runtime error %d", rc);
    else
        snprintf(reason, sizeof(reason), "Reason:\n\tstatus %d", rc);
    g_handler(rc, where, reason);
}

// d_nl -> the NEIGHBOR_LIST that made it
std::map<const void*, NEIGHBOR_LIST*>& list_registry() {
    static std::map<const void*, NEIGHBOR_LIST*> r;
    return r;
}

sponge_nl* list_of(const ATOM_GROUP* d_nl, const char* where) {
    NEIGHBOR_LIST* nl = NEIGHBOR_LIST::From_Device_List(d_nl);
    if (!nl || !nl->handle) {
        g_handler(SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND, where,
                  "Reason:\n\tthe neighbor list is not a d_nl of this library's NEIGHBOR_LIST");
        return nullptr;
    }
    return nl->handle;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) ++a;
    while (b > a && isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string unquote(std::string v) {
    v = trim(v);
    if (v.size() >= 2 && (v[0] == '"' || v[0] == '\'') && v.back() == v[0])
        return v.substr(1, v.size() - 2);
    return v;
}

// the text before a comment that is not inside quotes
std::string strip_comment(const std::string& line) {
    bool in_s = false, in_d = false;
    for (size_t i = 0; i < line.size(); ++i) {
        char c = line[i];
        if (c == '"' && !in_s) in_d = !in_d;
        else if (c == '\'' && !in_d) in_s = !in_s;
        else if (c == '#' && !in_s && !in_d) return line.substr(0, i);
    }
    return line;
}

}  // namespace

Boundary Boundary::Orthorhombic(float lx, float ly, float lz) {
    Boundary b;
    const float one = 1.0f;
    const float c[6] = {lx, 0.f, ly, 0.f, 0.f, lz};
    const float r[6] = {one / lx, 0.f, one / ly, 0.f, 0.f, one / lz};
    for (int k = 0; k < 6; ++k) {
        b.cell[k] = c[k];
        b.rcell[k] = r[k];
    }
    return b;
}

void Set_Error_Handler(ErrorHandler handler) { g_handler = handler ? handler : default_handler; }

void Throw_SPONGE_Error(int code, const char* where, const char* reason) {
    g_handler(code, where, reason);
}

// ------------------------------------------------------------------ CONTROLLER
// Legacy mdin lines are "key = value"; TOML tables "[table]" prefix their keys
// as "table_key" (SPONGE/control.cpp:87-94).  Keys are matched case-sensitively
// as written, like the reference's command map.
bool CONTROLLER::Load_Text(const std::string& text) {
    std::istringstream in(text);
    std::string line, table;
    while (std::getline(in, line)) {
        line = trim(strip_comment(line));
        if (line.empty()) continue;
        if (line.front() == '[') {
            size_t e = line.find(']');
            if (e == std::string::npos) return false;
            table = trim(line.substr(1, e - 1));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq));
        std::string val = unquote(line.substr(eq + 1));
        if (key.empty()) continue;
        commands[table.empty() ? key : table + "_" + key] = val;
    }
    return true;
}

bool CONTROLLER::Load(const std::string& path) {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    return Load_Text(ss.str());
}

bool CONTROLLER::Command_Exist(const char* key) const { return commands.count(key) != 0; }

bool CONTROLLER::Command_Exist(const char* prefix, const char* key) const {
    if (!prefix || !prefix[0]) return Command_Exist(key);
    return commands.count(std::string(prefix) + "_" + key) != 0;
}

const char* CONTROLLER::Command(const char* key) const {
    auto it = commands.find(key);
    return it == commands.end() ? nullptr : it->second.c_str();
}

const char* CONTROLLER::Command(const char* prefix, const char* key) const {
    if (!prefix || !prefix[0]) return Command(key);
    auto it = commands.find(std::string(prefix) + "_" + key);
    return it == commands.end() ? nullptr : it->second.c_str();
}

int CONTROLLER::Get_Int(const char* prefix, const char* key, int fallback) const {
    const char* v = Command(prefix, key);
    if (!v) return fallback;
    char* end = nullptr;
    long x = strtol(v, &end, 10);
    if (end == v || *end != 0) {
        std::string where = std::string(prefix ? prefix : "") + "_" + key;
        Throw_SPONGE_Error(SPONGE_AMD_ERROR_TYPE_ERROR_COMMAND, where.c_str(),
                           "Reason:\n\tthe value is not an integer");
        return fallback;
    }
    return (int)x;
}

float CONTROLLER::Get_Float(const char* prefix, const char* key, float fallback) const {
    const char* v = Command(prefix, key);
    if (!v) return fallback;
    char* end = nullptr;
    double x = strtod(v, &end);
    if (end == v || *end != 0) {
        std::string where = std::string(prefix ? prefix : "") + "_" + key;
        Throw_SPONGE_Error(SPONGE_AMD_ERROR_TYPE_ERROR_COMMAND, where.c_str(),
                           "Reason:\n\tthe value is not a float");
        return fallback;
    }
    return (float)x;
}

bool CONTROLLER::Get_Bool(const char* prefix, const char* key, bool fallback) const {
    const char* v = Command(prefix, key);
    if (!v) return fallback;
    std::string s(v);
    for (auto& c : s) c = (char)tolower((unsigned char)c);
    if (s == "1" || s == "true" || s == "yes" || s == "on") return true;
    if (s == "0" || s == "false" || s == "no" || s == "off") return false;
    std::string where = std::string(prefix ? prefix : "") + "_" + key;
    Throw_SPONGE_Error(SPONGE_AMD_ERROR_TYPE_ERROR_COMMAND, where.c_str(),
                       "Reason:\n\tthe value is not a bool");
    return fallback;
}

// ---------------------------------------------------------------- NEIGHBOR_LIST
void NEIGHBOR_LIST::Initial(CONTROLLER* controller, int atom_numbers_, float cutoff_, float skin_,
                            const Boundary boundary, void* stream_) {
    sponge_nl_options o;
    sponge_nl_default_options(&o);
    if (controller) {
        // [neighbor_list] keys (neighbor_list.cpp:276-300, 888-940)
        o.refresh_interval = controller->Get_Int("neighbor_list", "refresh_interval", o.refresh_interval);
        o.skin_permit = controller->Get_Float("neighbor_list", "skin_permit", o.skin_permit);
        o.max_neighbor_numbers =
            controller->Get_Int("neighbor_list", "max_neighbor_numbers", o.max_neighbor_numbers);
        o.max_atom_in_grid_numbers = controller->Get_Int(
            "neighbor_list", "max_atom_in_grid_numbers", o.max_atom_in_grid_numbers);
        o.check_overflow_interval = controller->Get_Int("neighbor_list", "check_overflow_interval",
                                                        o.check_overflow_interval);
        o.throw_error_when_overflow =
            controller->Get_Bool("neighbor_list", "throw_error_when_overflow", false) ? 1 : 0;
    }
    Initial_With_Options(o, atom_numbers_, cutoff_, skin_, boundary, stream_);
}

void NEIGHBOR_LIST::Initial_With_Options(const sponge_nl_options& o, int atom_numbers_,
                                         float cutoff_, float skin_, const Boundary boundary,
                                         void* stream_) {
    atom_numbers = atom_numbers_;
    cutoff = cutoff_;
    skin = skin_;
    stream = stream_;
    boundary0 = boundary;
    max_neighbor_numbers = o.max_neighbor_numbers;
    max_atom_in_grid_numbers = o.max_atom_in_grid_numbers;
    check_overflow_interval = o.check_overflow_interval;
    throw_error_when_overflow = o.throw_error_when_overflow != 0;
    check(sponge_nl_initial(&handle, atom_numbers, cutoff, skin, boundary.cell, boundary.rcell, &o,
                            stream),
          "NEIGHBOR_LIST::Initial");
    is_initialized = handle != nullptr;
    if (is_initialized) {
        void* a = nullptr;
        check(sponge_nl_atom_group_address(handle, &a), "NEIGHBOR_LIST::Initial");
        d_nl = static_cast<ATOM_GROUP*>(a);
        if (d_nl) list_registry()[d_nl] = this;
    }
}

NEIGHBOR_LIST* NEIGHBOR_LIST::From_Device_List(const ATOM_GROUP* d_nl) {
    auto& r = list_registry();
    auto it = r.find(d_nl);
    return it == r.end() ? nullptr : it->second;
}

void NEIGHBOR_LIST::Update(int* atom_local, int local_atom_numbers, int ghost_numbers, float* crd,
                           const Boundary boundary, int step, int update, int* excluded_list_start,
                           int* excluded_list, int* excluded_numbers) {
    if (!is_initialized) return;
    check(sponge_nl_update(handle, atom_local, local_atom_numbers, ghost_numbers, crd, boundary.cell,
                           boundary.rcell, step, update, excluded_list_start, excluded_list,
                           excluded_numbers),
          "NEIGHBOR_LIST::Update");
}

void NEIGHBOR_LIST::Check_Overflow(CONTROLLER* controller, int steps) {
    if (!is_initialized || check_overflow_interval <= 0 || (steps + 1) % check_overflow_interval != 0)
        return;
    int flags = 0;
    const int rc = sponge_nl_check_overflow(handle, &flags);
    if (rc == SPONGE_AMD_ERROR_OVERFLOW) {
        Throw_SPONGE_Error(SPONGE_AMD_ERROR_OVERFLOW, "NEIGHBOR_LIST::Check_Overflow",
                           "Reason:\n\tOverflow occured in neighbor searching");
        return;
    }
    if (rc) {
        check(rc, "NEIGHBOR_LIST::Check_Overflow");
        return;
    }
    if (!flags) return;
    // the reference's recovery: the module re-initialised with 1.1x the
    // limit that overflowed, the new limit written back to the commands
    sponge_nl_options o;
    check(sponge_nl_get_options(handle, &o), "NEIGHBOR_LIST::Check_Overflow");
    auto grow = [&](int& v, const char* key) {
        v = (int)(v * 1.1);
        if (controller) controller->commands[key] = std::to_string(v);
        printf("Overflow occured in neighbor searching.\nSPONGE will re-initialize the neighbor "
               "list module with %s = %d\n\n", key, v);
    };
    if (flags & SPONGE_AMD_NL_GRID_OVERFLOW)
        grow(o.max_atom_in_grid_numbers, "neighbor_list_max_atom_in_grid_numbers");
    if (flags & SPONGE_AMD_NL_LIST_OVERFLOW)
        grow(o.max_neighbor_numbers, "neighbor_list_max_neighbor_numbers");
    const Boundary b = boundary0;
    Clear();
    Initial_With_Options(o, atom_numbers, cutoff, skin, b, stream);
}

ATOM_GROUP* NEIGHBOR_LIST::Device_List() {
    void* d = nullptr;
    if (is_initialized) check(sponge_nl_atom_group(handle, &d), "NEIGHBOR_LIST::Device_List");
    return static_cast<ATOM_GROUP*>(d);
}

void NEIGHBOR_LIST::Clear() {
    if (d_nl) list_registry().erase(d_nl);
    d_nl = nullptr;
    if (handle) sponge_nl_clear(handle);
    handle = nullptr;
    is_initialized = false;
}

// ---------------------------------------------------- LENNARD_JONES_INFORMATION
void LENNARD_JONES_INFORMATION::Initial(CONTROLLER* controller, float cutoff_, int atom_numbers_,
                                        int atom_type_numbers_, const float* LJ_A,
                                        const float* LJ_B, const int* atom_LJ_type,
                                        void* stream_) {
    (void)controller;
    atom_numbers = atom_numbers_;
    atom_type_numbers = atom_type_numbers_;
    cutoff = cutoff_;
    stream = stream_;
    const int P = atom_type_numbers * (atom_type_numbers + 1) / 2;
    h_LJ_A.assign(LJ_A, LJ_A + P);
    h_LJ_B.assign(LJ_B, LJ_B + P);
    h_atom_LJ_type.assign(atom_LJ_type, atom_LJ_type + atom_numbers);
    check(sponge_lj_initial(&handle, atom_numbers, atom_type_numbers, LJ_A, LJ_B, atom_LJ_type,
                            cutoff, stream),
          "LENNARD_JONES_INFORMATION::Initial");
    if (handle) {
        check(sponge_lj_long_range_factor(handle, &long_range_factor),
              "LENNARD_JONES_INFORMATION::Initial");
        is_initialized = 1;
    }
}

void LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
    int atom_numbers_, int local_atom_numbers, int solvent_numbers, int ghost_numbers,
    const float* crd, const float* charge, float* frc, const Boundary boundary,
    const ATOM_GROUP* nl, float pme_beta, int need_atom_energy, float* atom_energy,
    int need_virial, float* atom_virial, float* atom_direct_pme_energy) {
    static const char* where =
        "LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial";
    if (!is_initialized) return;
    sponge_nl* list = list_of(nl, where);
    if (!list) return;
    check(sponge_lj_pme_direct_force(handle, list, atom_numbers_, local_atom_numbers,
                                     solvent_numbers, ghost_numbers, crd, charge, frc,
                                     boundary.cell, boundary.rcell, pme_beta, need_atom_energy,
                                     atom_energy, need_virial, atom_virial, atom_direct_pme_energy),
          where);
}

void LENNARD_JONES_INFORMATION::Long_Range_Correction(int need_pressure, float* d_virial,
                                                      int need_potential, float* d_potential,
                                                      float volume) {
    if (!is_initialized) return;
    check(sponge_lj_long_range_correction(handle, need_pressure, d_virial, need_potential,
                                          d_potential, volume),
          "LENNARD_JONES_INFORMATION::Long_Range_Correction");
}

void LENNARD_JONES_INFORMATION::Step_Print(CONTROLLER* controller) {
    (void)controller;
    if (!is_initialized) return;
    check(sponge_lj_step_print(handle, &h_LJ_energy_sum, &h_LJ_long_energy),
          "LENNARD_JONES_INFORMATION::Step_Print");
}

void LENNARD_JONES_INFORMATION::Clear() {
    if (handle) sponge_lj_clear(handle);
    handle = nullptr;
    is_initialized = 0;
}

// ------------------------------------------------------- SOLVENT_LENNARD_JONES
void SOLVENT_LENNARD_JONES::Initial(CONTROLLER* controller, LENNARD_JONES_INFORMATION* lj,
                                    void* lj_soft, int atom_numbers, int residue_numbers,
                                    const int* res_start, const float* mass, bool default_enable,
                                    void* stream_) {
    (void)lj_soft;
    if (!lj || !lj->is_initialized) return;
    // solvent_LJ = true / false; absent -> -1 (the default gating)
    int key = -1;
    if (controller && controller->Command_Exist("solvent_LJ"))
        key = controller->Get_Bool("", "solvent_LJ", false) ? 1 : 0;
    stream = stream_;
    lj_info = lj;
    check(sponge_solvent_lj_initial(&handle, lj->handle, atom_numbers, residue_numbers, res_start,
                                    mass, key, default_enable ? 1 : 0, stream),
          "SOLVENT_LENNARD_JONES::Initial");
    if (!handle) return;
    int enabled = 0;
    check(sponge_solvent_lj_info(handle, &enabled, &solvent_numbers, &solvent_start,
                                 &water_points),
          "SOLVENT_LENNARD_JONES::Initial");
    is_initialized = enabled;
    if (!enabled) solvent_numbers = 0;
    local_solvent_numbers = solvent_numbers;
}

void SOLVENT_LENNARD_JONES::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
    int atom_numbers, int residue_numbers, const int* d_res_start, const float* crd,
    const float* charge, float* frc, const Boundary boundary, const ATOM_GROUP* nl,
    float pme_beta, int need_atom_energy, float* atom_energy, int need_virial,
    float* atom_lj_virial, float* atom_direct_pme_energy) {
    static const char* where =
        "SOLVENT_LENNARD_JONES::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial";
    (void)d_res_start;
    if (!is_initialized) return;
    sponge_nl* list = list_of(nl, where);
    if (!list) return;
    check(sponge_solvent_lj_direct_force(handle, list, atom_numbers, residue_numbers, crd, charge,
                                         frc, boundary.cell, boundary.rcell, pme_beta,
                                         need_atom_energy, atom_energy, need_virial,
                                         atom_lj_virial, atom_direct_pme_energy),
          where);
}

void SOLVENT_LENNARD_JONES::Clear() {
    if (handle) sponge_solvent_lj_clear(handle);
    handle = nullptr;
    is_initialized = 0;
    solvent_numbers = local_solvent_numbers = 0;
}

// --------------------------------------------------------------- Particle_Mesh
void Particle_Mesh::Initial(CONTROLLER* controller, int atom_numbers_, const Boundary boundary,
                            float cutoff_, void* stream_) {
    sponge_pme_options o;
    sponge_pme_default_options(&o);
    if (controller) {
        // [PM] keys (PM_force.cpp:315-398) and [PME] keys (:417-502)
        o.fftx = controller->Get_Int("PM", "fftx", o.fftx);
        o.ffty = controller->Get_Int("PM", "ffty", o.ffty);
        o.fftz = controller->Get_Int("PM", "fftz", o.fftz);
        o.grid_spacing = controller->Get_Float("PM", "grid_spacing", o.grid_spacing);
        o.direct_tolerance = controller->Get_Float("PM", "Direct_Tolerance", o.direct_tolerance);
        o.update_interval = controller->Get_Int("PME", "update_interval", o.update_interval);
        o.calculate_reciprocal_part =
            controller->Get_Bool("PME", "calculate_reciprocal_part", true) ? 1 : 0;
        o.calculate_excluded_part =
            controller->Get_Bool("PME", "calculate_excluded_part", true) ? 1 : 0;
        if (controller->Get_Bool("PME", "replaced_by_PMC_IZ", false))
            Throw_SPONGE_Error(SPONGE_AMD_ERROR_NOT_IMPLEMENTED, "Particle_Mesh::Initial",
                               "Reason:\n\tPMC-IZ is not part of this path");
        if (controller->Get_Int("PM", "MPI_size", 0) > 1)
            Throw_SPONGE_Error(SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND, "Particle_Mesh::Initial",
                               "Reason:\n\t Multi-process PME is not supported yet.");
    }
    atom_numbers = atom_numbers_;
    cutoff = cutoff_;
    stream = stream_;
    update_interval = o.update_interval;
    calculate_reciprocal_part = o.calculate_reciprocal_part != 0;
    calculate_excluded_part = o.calculate_excluded_part != 0;
    check(sponge_pme_initial(&handle, atom_numbers, boundary.cell, boundary.rcell, cutoff, &o, stream),
          "Particle_Mesh::Initial");
    if (handle) {
        int f[3];
        check(sponge_pme_info(handle, &beta, f), "Particle_Mesh::Initial");
        fftx = f[0];
        ffty = f[1];
        fftz = f[2];
        d_direct_atom_energy = sponge_pme_direct_atom_energy(handle);
        is_initialized = 1;
        for (int k = 0; k < 3; ++k) min_corner[k] = 0.0f;
        max_corner[0] = boundary.cell[0];
        max_corner[1] = boundary.cell[2];
        max_corner[2] = boundary.cell[5];
    }
}

void Particle_Mesh::Update_Box(const Boundary boundary, LTMatrix3 g, float dt) {
    if (!is_initialized) return;
    check(sponge_pme_update_box(handle, boundary.cell, boundary.rcell), "Particle_Mesh::Update_Box");
    auto move = [&](float* r) {
        const float x = r[0] + dt * (r[0] * g.a11 + r[1] * g.a21 + r[2] * g.a31);
        const float y = r[1] + dt * (r[1] * g.a22 + r[2] * g.a32);
        const float z = r[2] + dt * r[2] * g.a33;
        r[0] = x;
        r[1] = y;
        r[2] = z;
    };
    move(min_corner);
    move(max_corner);
}

void Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial(
    const float* crd, const Boundary boundary, const float* charge, float* force, int need_virial,
    int need_energy, float* d_virial, float* d_potential, int step) {
    if (!is_initialized) return;
    check(sponge_pme_reciprocal_force(handle, crd, boundary.cell, boundary.rcell, charge, force,
                                      need_virial, need_energy, d_virial, d_potential, step),
          "Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial");
}

void Particle_Mesh::MPI_PME_Excluded_Force_With_Atom_Energy(
    int local_atom_numbers, const int* atom_local, const int* atom_local_id, const float* crd,
    const Boundary boundary, const float* charge, const int* excluded_list_start,
    const int* excluded_list, const int* excluded_numbers, float* frc, int need_energy,
    float* atom_energy, int need_virial, float* atom_virial) {
    if (!is_initialized) return;
    check(sponge_pme_excluded_force(handle, local_atom_numbers, atom_local, atom_local_id, crd,
                                    boundary.cell, boundary.rcell, charge, excluded_list_start,
                                    excluded_list, excluded_numbers, frc, need_energy, atom_energy,
                                    need_virial, atom_virial),
          "Particle_Mesh::MPI_PME_Excluded_Force_With_Atom_Energy");
}

void Particle_Mesh::Step_Print(CONTROLLER* controller) {
    (void)controller;
    if (!is_initialized) return;
    check(sponge_pme_step_print(handle, &ee_ene, &direct_ene, &reciprocal_ene, &self_ene,
                                &correction_ene),
          "Particle_Mesh::Step_Print");
}

void Particle_Mesh::Clear() {
    if (handle) sponge_pme_clear(handle);
    handle = nullptr;
    is_initialized = 0;
}

}  // namespace sponge_amd
