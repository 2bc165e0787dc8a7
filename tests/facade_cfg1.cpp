// Drives the C++ module facades in the order of SPONGE/main.cpp's force
// evaluation (neighbor_list.Update, lj / solvent_lj direct calls with
// neighbor_list.d_nl, Long_Range_Correction, the PME excluded and reciprocal
// calls, Step_Print; main.cpp:1355-1626) on one system read from a folder,
// then a box change through Particle_Mesh::Update_Box, and writes every
// result to the same folder.  Test program: built and run by
// tests/test_gpu_facade.py.
//
// Folder contents (raw little-endian arrays): meta.txt ("N T R E cutoff skin
// lx ly lz scale dt"), mdin.txt, crd, charge, mass (float32), type (int32),
// A, B (float32), es, en, el, res (int32), crd2 (float32, the coordinates in
// the box scaled by `scale`).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sponge_amd/host/sponge_modules.hpp"

using namespace sponge_amd;

static std::string dir;

template <typename T>
static std::vector<T> load(const char* name, size_t n) {
    std::vector<T> v(n ? n : 1);
    std::ifstream f(dir + "/" + name, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("missing ") + name);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
    if ((size_t)f.gcount() != n * sizeof(T)) throw std::runtime_error(std::string("short ") + name);
    return v;
}

template <typename T>
static void save(const char* name, const T* p, size_t n) {
    std::ofstream f(dir + "/" + name, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

static void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
static T* dev(const std::vector<T>& h) {
    T* d = nullptr;
    hip_ok(hipMalloc(&d, h.size() * sizeof(T)), "hipMalloc");
    hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

template <typename T>
static std::vector<T> host(const T* d, size_t n) {
    std::vector<T> h(n);
    hip_ok(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

static void thrower(int code, const char* where, const char* reason) {
    throw std::runtime_error(std::string(where) + " " + std::to_string(code) + " " +
                             (reason ? reason : ""));
}

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s <folder>\n", argv[0]);
        return 2;
    }
    dir = argv[1];
    Set_Error_Handler(thrower);
    try {
        int N, T, R, E;
        float cutoff, skin, lx, ly, lz, scale, dt;
        {
            std::ifstream m(dir + "/meta.txt");
            m >> N >> T >> R >> E >> cutoff >> skin >> lx >> ly >> lz >> scale >> dt;
            if (!m) throw std::runtime_error("meta.txt");
        }
        const int P = T * (T + 1) / 2;
        auto crd = load<float>("crd", 3 * (size_t)N), charge = load<float>("charge", N),
             mass = load<float>("mass", N), A = load<float>("A", P), B = load<float>("B", P),
             crd2 = load<float>("crd2", 3 * (size_t)N);
        auto type = load<int>("type", N), es = load<int>("es", N), en = load<int>("en", N),
             el = load<int>("el", E), res = load<int>("res", R);

        CONTROLLER controller;
        if (!controller.Load(dir + "/mdin.txt")) throw std::runtime_error("mdin.txt");
        const Boundary box = Boundary::Orthorhombic(lx, ly, lz);

        float *d_crd = dev(crd), *d_charge = dev(charge);
        int *d_es = dev(es), *d_en = dev(en), *d_el = dev(el);
        float* d_frc = dev(std::vector<float>(3 * (size_t)N, 0.f));
        float* d_energy = dev(std::vector<float>(N, 0.f));
        float* d_atom_virial = dev(std::vector<float>(6 * (size_t)N, 0.f));
        float* d_virial = dev(std::vector<float>(6, 0.f));
        float* d_potential = dev(std::vector<float>(1, 0.f));

        NEIGHBOR_LIST neighbor_list;
        neighbor_list.Initial(&controller, N, cutoff, skin, box);
        LENNARD_JONES_INFORMATION lj;
        lj.Initial(&controller, cutoff, N, T, A.data(), B.data(), type.data());
        SOLVENT_LENNARD_JONES solvent_lj;
        solvent_lj.Initial(&controller, &lj, nullptr, N, R, res.data(), mass.data(), true);
        Particle_Mesh pm;
        pm.Initial(&controller, N, box, cutoff);

        // one energy step of the force loop
        neighbor_list.Update(nullptr, N, 0, d_crd, box, 0,
                             NEIGHBOR_LIST::FORCED_UPDATE, d_es, d_el, d_en);
        lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
            N, N, solvent_lj.local_solvent_numbers, 0, d_crd, d_charge, d_frc, box,
            neighbor_list.d_nl, pm.beta, 1, d_energy, 1, d_atom_virial, pm.d_direct_atom_energy);
        solvent_lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
            N, R, nullptr, d_crd, d_charge, d_frc, box, neighbor_list.d_nl, pm.beta, 1, d_energy,
            1, d_atom_virial, pm.d_direct_atom_energy);
        lj.Long_Range_Correction(1, d_virial, 1, d_potential, lx * ly * lz);
        pm.MPI_PME_Excluded_Force_With_Atom_Energy(N, nullptr, nullptr, d_crd, box, d_charge, d_es,
                                                   d_el, d_en, d_frc, 1, d_energy, 1,
                                                   d_atom_virial);
        pm.PME_Reciprocal_Force_With_Energy_And_Virial(d_crd, box, d_charge, d_frc, 1, 1,
                                                       d_virial, d_potential, 0);
        hip_ok(hipDeviceSynchronize(), "step");
        lj.Step_Print(&controller);
        pm.Step_Print(&controller);

        save("out_frc", host(d_frc, 3 * (size_t)N).data(), 3 * (size_t)N);
        save("out_energy", host(d_energy, N).data(), N);
        save("out_atom_virial", host(d_atom_virial, 6 * (size_t)N).data(), 6 * (size_t)N);
        save("out_virial", host(d_virial, 6).data(), 6);
        save("out_potential", host(d_potential, 1).data(), 1);
        save("out_direct_atom_energy", host(pm.d_direct_atom_energy, N).data(), N);
        const float scalars[] = {lj.h_LJ_energy_sum, lj.h_LJ_long_energy, pm.ee_ene,
                                 pm.direct_ene, pm.reciprocal_ene, pm.self_ene,
                                 pm.correction_ene, pm.beta, (float)pm.fftx, (float)pm.ffty,
                                 (float)pm.fftz, (float)solvent_lj.solvent_numbers,
                                 (float)solvent_lj.is_initialized};
        save("out_scalars", scalars, sizeof(scalars) / sizeof(float));

        // the list's rows through d_nl, as a module that reads it directly sees them
        if (neighbor_list.Device_List() != neighbor_list.d_nl)
            throw std::runtime_error("Device_List() is not d_nl");
        const int M = neighbor_list.max_neighbor_numbers;
        auto groups = host(neighbor_list.d_nl, N);
        std::vector<int> counts(N), serial((size_t)N * M, -1);
        for (int i = 0; i < N; ++i) {
            counts[i] = groups[i].atom_numbers;
            if (counts[i] < 0 || counts[i] > M) throw std::runtime_error("row length");
            if (counts[i]) {
                auto row = host(groups[i].atom_serial, counts[i]);
                for (int k = 0; k < counts[i]; ++k) serial[(size_t)i * M + k] = row[k];
            }
        }
        save("out_nl_counts", counts.data(), N);
        save("out_nl_serial", serial.data(), serial.size());

        // a list pointer that no NEIGHBOR_LIST made is refused
        int refused = 0;
        try {
            lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
                N, N, 0, 0, d_crd, d_charge, d_frc, box,
                reinterpret_cast<const ATOM_GROUP*>(d_crd), pm.beta, 0, nullptr, 0, nullptr,
                nullptr);
        } catch (const std::runtime_error&) {
            refused = 1;
        }

        // box change: the box scaled by `scale` over one step dt at the rate g
        const float rate = (scale - 1.0f) / dt;
        const LTMatrix3 g = {rate, 0.f, rate, 0.f, 0.f, rate};
        const Boundary box2 = Boundary::Orthorhombic(lx * scale, ly * scale, lz * scale);
        pm.Update_Box(box2, g, dt);
        float* d_crd2 = dev(crd2);
        float* d_frc2 = dev(std::vector<float>(3 * (size_t)N, 0.f));
        float* d_virial2 = dev(std::vector<float>(6, 0.f));
        float* d_potential2 = dev(std::vector<float>(1, 0.f));
        pm.PME_Reciprocal_Force_With_Energy_And_Virial(d_crd2, box2, d_charge, d_frc2, 1, 1,
                                                       d_virial2, d_potential2, 0);
        hip_ok(hipDeviceSynchronize(), "box change");
        pm.Step_Print(&controller);
        save("out_frc2", host(d_frc2, 3 * (size_t)N).data(), 3 * (size_t)N);
        save("out_virial2", host(d_virial2, 6).data(), 6);
        save("out_potential2", host(d_potential2, 1).data(), 1);
        const float scalars2[] = {pm.reciprocal_ene, pm.self_ene,
                                  pm.min_corner[0], pm.min_corner[1], pm.min_corner[2],
                                  pm.max_corner[0], pm.max_corner[1], pm.max_corner[2],
                                  (float)refused};
        save("out_scalars2", scalars2, sizeof(scalars2) / sizeof(float));

        solvent_lj.Clear();
        lj.Clear();
        pm.Clear();
        neighbor_list.Clear();
        for (void* p : {(void*)d_crd, (void*)d_charge, (void*)d_es, (void*)d_en, (void*)d_el,
                        (void*)d_frc, (void*)d_energy, (void*)d_atom_virial, (void*)d_virial,
                        (void*)d_potential, (void*)d_crd2, (void*)d_frc2, (void*)d_virial2,
                        (void*)d_potential2})
            hip_ok(hipFree(p), "hipFree");
    } catch (const std::exception& e) {
        fprintf(stderr, "facade_cfg1: %s\n", e.what());
        return 1;
    }
    printf("done\n");
    return 0;
}
