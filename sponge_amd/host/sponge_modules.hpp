// C++ facades with the reference's module interfaces over the C ABI of
// include/sponge_amd.h, so SPONGE's step loop (SPONGE/main.cpp:1355-1626) can
// call the MI355X path where it calls its own modules.
//
//   NEIGHBOR_LIST              SPONGE/neighbor_list/neighbor_list.h:7-127
//   LENNARD_JONES_INFORMATION  SPONGE/Lennard_Jones_force/Lennard_Jones_force.h:87-173
//   SOLVENT_LENNARD_JONES      SPONGE/Lennard_Jones_force/solvent_LJ.h:9-43
//   Particle_Mesh              SPONGE/PM_force/PM_force.h:152-209
//
// The force calls take the list as the reference's callers pass it,
// NEIGHBOR_LIST::d_nl (SPONGE/main.cpp:1452-1470): a d_nl made by a
// NEIGHBOR_LIST of this library names that list, whose cell-ordered form the
// kernels read; any other pointer is an error.
//
// Configuration comes from the same mdin keys, legacy "key = value" or TOML
// tables flattened to "<table>_<key>" (SPONGE/control.cpp:50-94).  Every
// non-zero status of the C ABI goes to an error handler that, by default,
// prints and exits with the code, as Throw_SPONGE_Error does
// (SPONGE/utils/control/error.hpp:31-147); tests install a throwing one.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../include/sponge_amd.h"

namespace sponge_amd {

// Lower-triangular 3x3 matrix, the reference's element order (SPONGE/utils/vector.hpp:257-259).
struct LTMatrix3 {
    float a11, a21, a22, a31, a32, a33;
};

// One atom's neighbor row, the reference's layout (SPONGE/common.h:177-182).
struct ATOM_GROUP {
    int atom_numbers;
    int ghost_numbers;
    int* atom_serial;
};

// The reference's Boundary: cell and rcell LTMatrix3 (SPONGE/utils/boundary.h:16-20).
struct Boundary {
    float cell[6];
    float rcell[6];
    static Boundary Orthorhombic(float lx, float ly, float lz);
};

using ErrorHandler = void (*)(int code, const char* where, const char* reason);
void Set_Error_Handler(ErrorHandler handler);
void Throw_SPONGE_Error(int code, const char* where, const char* reason);

// mdin command map (the subset of CONTROLLER the modules read).
class CONTROLLER {
   public:
    std::map<std::string, std::string> commands;
    // parses a legacy or TOML mdin file; returns false when it cannot be read
    bool Load(const std::string& path);
    bool Load_Text(const std::string& text);
    bool Command_Exist(const char* key) const;
    bool Command_Exist(const char* prefix, const char* key) const;
    const char* Command(const char* key) const;
    const char* Command(const char* prefix, const char* key) const;
    int Get_Int(const char* prefix, const char* key, int fallback) const;
    float Get_Float(const char* prefix, const char* key, float fallback) const;
    bool Get_Bool(const char* prefix, const char* key, bool fallback) const;
};

struct NEIGHBOR_LIST {
    enum NEIGHBOR_LIST_UPDATE_PARAMETER { CONDITIONAL_UPDATE = 0, FORCED_UPDATE = 1 };
    sponge_nl* handle = nullptr;
    bool is_initialized = false;
    float cutoff = 0, skin = 0;
    int atom_numbers = 0;
    int max_neighbor_numbers = 1200;
    int max_atom_in_grid_numbers = 150;
    int check_overflow_interval = 150;
    bool throw_error_when_overflow = false;
    void* stream = nullptr;
    Boundary boundary0{};  // the box of Initial (a re-initialisation reuses it)
    // the name of this list for the force calls (neighbor_list.h:42); a device
    // array of ATOM_GROUP whose rows Device_List() fills.  A re-initialisation
    // (Check_Overflow) gives a new value, as the reference's does.
    ATOM_GROUP* d_nl = nullptr;

    // NEIGHBOR_LIST::Initial (neighbor_list.cpp:871-972)
    void Initial(CONTROLLER* controller, int atom_numbers, float cutoff, float skin,
                 const Boundary boundary, void* stream = nullptr);
    // NEIGHBOR_LIST::Update (neighbor_list.cpp:974-1025)
    void Update(int* atom_local, int local_atom_numbers, int ghost_numbers, float* crd,
                const Boundary boundary, int step, int update, int* excluded_list_start = nullptr,
                int* excluded_list = nullptr, int* excluded_numbers = nullptr);
    // NEIGHBOR_LIST::Check_Overflow (neighbor_list.cpp:1027-1103): every
    // check_overflow_interval steps; on an overflow the module is
    // re-initialised with 1.1x the limit that overflowed (the new value also
    // written to the controller's commands), or, with
    // throw_error_when_overflow, the error handler gets spongeErrorOverflow
    void Check_Overflow(CONTROLLER* controller, int steps);
    // fills d_nl's rows (the reference's half list, max_neighbor_numbers per
    // atom) for modules that read the list directly; returns d_nl
    ATOM_GROUP* Device_List();
    void Clear();
    // the list that d_nl names, or null
    static NEIGHBOR_LIST* From_Device_List(const ATOM_GROUP* d_nl);

   private:
    void Initial_With_Options(const sponge_nl_options& o, int atom_numbers, float cutoff,
                              float skin, const Boundary boundary, void* stream);
};

struct LENNARD_JONES_INFORMATION {
    sponge_lj* handle = nullptr;
    int is_initialized = 0;
    int atom_numbers = 0;
    int atom_type_numbers = 0;
    float cutoff = 10.0f;
    float long_range_factor = 0;
    float h_LJ_energy_sum = 0, h_LJ_long_energy = 0;
    std::vector<float> h_LJ_A, h_LJ_B;
    std::vector<int> h_atom_LJ_type;
    void* stream = nullptr;

    // Initial (Lennard_Jones_force.cpp:147-250) from host tables in memory
    // convention (A x12, B x6 of the file values) -- the arrays the reference
    // loads from LJ_in_file (SPONGE/xponge/load/native/lj.hpp:8-69)
    void Initial(CONTROLLER* controller, float cutoff, int atom_numbers, int atom_type_numbers,
                 const float* LJ_A, const float* LJ_B, const int* atom_LJ_type,
                 void* stream = nullptr);
    // LJ_PME_Direct_Force_With_Atom_Energy_And_Virial (:326-384); nl is the
    // list's d_nl
    void LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
        int atom_numbers, int local_atom_numbers, int solvent_numbers, int ghost_numbers,
        const float* crd, const float* charge, float* frc, const Boundary boundary,
        const ATOM_GROUP* nl, float pme_beta, int need_atom_energy, float* atom_energy,
        int need_virial, float* atom_virial, float* atom_direct_pme_energy);
    // Long_Range_Correction (:287-308)
    void Long_Range_Correction(int need_pressure, float* d_virial, int need_potential,
                               float* d_potential, float volume);
    // Step_Print (:386-400): LJ_short, LJ_long
    void Step_Print(CONTROLLER* controller);
    void Clear();
};

// solvent_LJ (solvent_LJ.h:9-43).  MD_INFORMATION's update groups and masses
// come in as host arrays: residue_numbers residues of contiguous atoms from
// res_start[r], and the atoms' masses.
struct SOLVENT_LENNARD_JONES {
    sponge_solvent_lj* handle = nullptr;
    int is_initialized = 0;
    int solvent_numbers = 0;
    int solvent_start = 0;  // first residue of the water tail
    int water_points = 0;
    int local_solvent_numbers = 0;  // one domain: solvent_numbers
    LENNARD_JONES_INFORMATION* lj_info = nullptr;
    void* stream = nullptr;

    // Initial (solvent_LJ.cpp:149-248): the mdin key solvent_LJ, else on when
    // default_enable and there are at least 10 residues (main.cpp:1303-1307
    // passes mode >= NVT); lj_soft is accepted and unused (no soft-core tail)
    void Initial(CONTROLLER* controller, LENNARD_JONES_INFORMATION* lj, void* lj_soft,
                 int atom_numbers, int residue_numbers, const int* res_start, const float* mass,
                 bool default_enable, void* stream = nullptr);
    // LJ_PME_Direct_Force_With_Atom_Energy_And_Virial (solvent_LJ.cpp:256-432);
    // d_res_start is accepted for the reference's signature (the residues of
    // Initial are used); nl is the list's d_nl
    void LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
        int atom_numbers, int residue_numbers, const int* d_res_start, const float* crd,
        const float* charge, float* frc, const Boundary boundary, const ATOM_GROUP* nl,
        float pme_beta, int need_atom_energy, float* atom_energy, int need_virial,
        float* atom_lj_virial, float* atom_direct_pme_energy);
    void Clear();
};

struct Particle_Mesh {
    sponge_pme* handle = nullptr;
    int is_initialized = 0;
    int atom_numbers = 0;
    int fftx = 0, ffty = 0, fftz = 0;
    float beta = 0, cutoff = 0;
    int update_interval = 1;
    bool calculate_reciprocal_part = true, calculate_excluded_part = true;
    float direct_ene = 0, reciprocal_ene = 0, self_ene = 0, correction_ene = 0, ee_ene = 0;
    float* d_direct_atom_energy = nullptr;  // written by the LJ direct call
    void* stream = nullptr;

    // Initial (PM_force.cpp:295-654): [PM] fftx/ffty/fftz, grid_spacing,
    // Direct_Tolerance; [PME] update_interval, calculate_reciprocal_part,
    // calculate_excluded_part
    void Initial(CONTROLLER* controller, int atom_numbers, const Boundary boundary, float cutoff,
                 void* stream = nullptr);
    // PME_Reciprocal_Force_With_Energy_And_Virial (:1070-1168)
    void PME_Reciprocal_Force_With_Energy_And_Virial(const float* crd, const Boundary boundary,
                                                     const float* charge, float* force,
                                                     int need_virial, int need_energy,
                                                     float* d_virial, float* d_potential,
                                                     int step);
    // Update_Box (PM_force.cpp:1254-1266): BC / virial tables and the
    // neutralizing factor for the new box; the domain corners (one domain:
    // the origin and the box diagonal) moved by dt * (r g) as the
    // reference moves them (:1244-1252)
    void Update_Box(const Boundary boundary, LTMatrix3 g, float dt);
    float min_corner[3] = {0, 0, 0}, max_corner[3] = {0, 0, 0};
    // MPI_PME_Excluded_Force_With_Atom_Energy (:1819-1842)
    void MPI_PME_Excluded_Force_With_Atom_Energy(int local_atom_numbers, const int* atom_local,
                                                 const int* atom_local_id, const float* crd,
                                                 const Boundary boundary, const float* charge,
                                                 const int* excluded_list_start,
                                                 const int* excluded_list,
                                                 const int* excluded_numbers, float* frc,
                                                 int need_energy, float* atom_energy,
                                                 int need_virial, float* atom_virial);
    // Step_Print (:1844-1869): PM and its pieces
    void Step_Print(CONTROLLER* controller);
    void Clear();
};

}  // namespace sponge_amd
