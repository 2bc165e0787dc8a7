"""The C++ module facades compile against the library and read the
reference's mdin keys (legacy and TOML); no GPU calls."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include <stdexcept>
#include "sponge_amd/host/sponge_modules.hpp"
using namespace sponge_amd;
static void thrower(int code, const char* where, const char* reason) {
    throw std::runtime_error(std::string(where) + ":" + std::to_string(code));
}
int main() {
    Set_Error_Handler(thrower);
    CONTROLLER c;
    c.Load_Text(
        "cutoff = 10.0   # top level\n"
        "skin = 2.0\n"
        "[neighbor_list]\n"
        "max_neighbor_numbers = 800\n"
        "skin_permit = 0.4\n"
        "[PM]\n"
        "fftx = 96\n"
        "Direct_Tolerance = 1e-6\n"
        "print_detail = true\n"
        "[PME]\n"
        "update_interval = 2\n"
        "mode = \"nve\"\n");
    int ok = 1;
    ok &= c.Command_Exist("cutoff") && c.Get_Float("", "cutoff", 0) == 10.0f;
    ok &= c.Get_Int("neighbor_list", "max_neighbor_numbers", 0) == 800;
    ok &= c.Get_Float("neighbor_list", "skin_permit", 0) == 0.4f;
    ok &= c.Get_Int("PM", "fftx", 0) == 96;
    ok &= c.Get_Float("PM", "Direct_Tolerance", 0) == 1e-6f;
    ok &= c.Get_Bool("PM", "print_detail", false);
    ok &= c.Get_Int("PME", "update_interval", 1) == 2;
    ok &= std::string(c.Command("PME", "mode")) == "nve";
    ok &= !c.Command_Exist("PM", "ffty") && c.Get_Int("PM", "ffty", -1) == -1;
    bool threw = false;
    try { c.Get_Int("PM", "Direct_Tolerance", 0); } catch (std::exception&) { threw = true; }
    ok &= threw;
    Boundary b = Boundary::Orthorhombic(10, 20, 30);
    ok &= b.rcell[0] == 0.1f && b.cell[5] == 30.0f;
    std::printf(ok ? "ok\n" : "FAIL\n");
    return ok ? 0 : 1;
}
'''


def test_facade_reads_mdin_keys(tmp_path):
    lib = os.path.join(ROOT, "sponge_amd", "libsponge_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-I", ROOT, str(src), lib, "-o", str(exe),
                    f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr


def test_facade_force_loop_program_builds(tmp_path):
    """The GPU facade test's program (tests/facade_cfg1.cpp: the reference's
    force loop through NEIGHBOR_LIST / LENNARD_JONES_INFORMATION /
    SOLVENT_LENNARD_JONES / Particle_Mesh with neighbor_list.d_nl) compiles
    and links against the library; it runs in tests/test_gpu_facade.py."""
    lib = os.path.join(ROOT, "sponge_amd", "libsponge_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    exe = tmp_path / "facade_cfg1"
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", ROOT,
                    "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "facade_cfg1.cpp"),
                    lib, "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{os.path.dirname(lib)}",
                    "-o", str(exe)], check=True)
    assert exe.exists()
