"""The drop-in shim (sm_2d_amd/csrc/sm2d_shim.cpp) compiled against the reference's own boundary
headers (src/include/cuda/gpu_transport_wrapper.hpp:21-95 and the grid, LUT and config headers it
uses) and linked against libsm2d_hip.so: the symbols the reference runner calls resolve with the
exact signatures of that header.  A small program takes the address of each entry point through
the reference's declarations and constructs and destroys a DeviceLUTWrapper (no GPU call).

Needs g++ and the SM_2D checkout (its headers are not part of this repository); skipped without
them."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = os.path.join(os.environ.get("SM2D_REFERENCE", "/root/reference"), "src", "include")
LIB_DIR = os.path.join(ROOT, "sm_2d_amd")

MAIN = r"""
#include <string>
#include <vector>
#include "cuda/gpu_transport_wrapper.hpp"

int main() {
    bool (*run)(float, float, float, float, float, float, float, float, int, unsigned int, int, int, float,
                float, float, float, int, int, const AngularGrid&, const EnergyGrid&, const ::DeviceRLUT&,
                const sm_2d::TransportConfig&, std::vector<std::vector<double>>&) =
        &sm_2d::run_k1k6_pipeline_transport;
    bool (*init)(const ::RLUT&, sm_2d::DeviceLUTWrapper&) = &sm_2d::init_device_lut;
    std::string (*name)() = &sm_2d::get_gpu_name_internal;
    sm_2d::DeviceLUTWrapper w;
    return (run && init && name && w.get() == nullptr) ? 0 : 1;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                    reason="needs the SM_2D checkout's headers and g++")
def test_shim_links_with_reference_signatures(tmp_path):
    lib = os.path.join(LIB_DIR, "libsm2d_hip.so")
    assert os.path.exists(lib), "build the library first (python -m sm_2d_amd.build)"
    main = tmp_path / "main.cpp"
    main.write_text(MAIN)
    inc = ["-I" + REF_INCLUDE, "-I" + os.path.join(ROOT, "include")]
    shim_o, main_o, exe = tmp_path / "shim.o", tmp_path / "main.o", tmp_path / "shim_check"
    for src, obj in ((os.path.join(ROOT, "sm_2d_amd", "csrc", "sm2d_shim.cpp"), shim_o), (str(main), main_o)):
        r = subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-fPIC", *inc, "-c", src, "-o", str(obj)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", str(shim_o), str(main_o), "-o", str(exe), "-L" + LIB_DIR, "-lsm2d_hip",
                        "-Wl,-rpath," + LIB_DIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the mangled names the reference runner's object file refers to
    syms = subprocess.run(["nm", "-C", str(shim_o)], capture_output=True, text=True).stdout
    for s in ("sm_2d::run_k1k6_pipeline_transport(float, float, float, float, float, float, float, float, int, "
              "unsigned int, int, int, float, float, float, float, int, int, AngularGrid const&, EnergyGrid const&, "
              "DeviceRLUT const&, sm_2d::TransportConfig const&, std::vector<std::vector<double, "
              "std::allocator<double> >, std::allocator<std::vector<double, std::allocator<double> > > >&)",
              "sm_2d::init_device_lut(RLUT const&, sm_2d::DeviceLUTWrapper&)",
              "sm_2d::get_gpu_name_internal[abi:cxx11]()",
              "sm_2d::DeviceLUTWrapper::DeviceLUTWrapper()",
              "sm_2d::DeviceLUTWrapper::~DeviceLUTWrapper()"):
        assert s in syms, s
    assert subprocess.run([str(exe)]).returncode == 0
