"""CPU: the launcher shim of INTEGRATION.md section 1 against the reference's own prototypes.

The shim (the reference-side binding a maintainer adds) is extracted from INTEGRATION.md and
compiled after the launcher declarations read from the reference's headers
(cpp_src/*/kernel_*.h); a second translation unit takes the address of every declared launcher
and is linked with the shim and libburyn_amd.so.  A shim function whose signature differs from
the reference's declaration leaves that declaration undefined and the link fails.  Needs the
reference checkout (development container only).
"""
import os
import re
import subprocess

import pytest

from helpers import REPO

REF = "/root/reference/cpp_src"
HEADERS = ["neuron_groups/kernel_neuron_group_evolve.h", "connections/kernel_connection_propagate.h",
           "stimuli/kernel_stimulus_apply.h", "plasticities/kernel_triplet_plasticity_update.h",
           "plasticities/kernel_istdp_plasticity_update.h", "plasticities/kernel_stdp_plasticity_update.h",
           "monitors/kernel_historicize_spikes.h", "monitors/kernel_historicize_membrane_potentials.h",
           "gpu_core/kernel_init_neuron_rng_states.h"]
LIB = os.path.join(REPO, "example-cuda-spiking-nn_amd", "lib")


def _shim_source() -> str:
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    section = text[text.index("## 1."):text.index("## 2.")]
    return re.search(r"```cpp\n(.*?)```", section, re.S).group(1)


def _reference_prototypes():
    protos = []
    for h in HEADERS:
        src = open(os.path.join(REF, h)).read()
        for m in re.finditer(r"void\s+(kernel_\w+_launch)\s*\(([^;{]*)\)\s*;", src):
            protos.append((m.group(1), " ".join(m.group(0).split())))
    return protos


@pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference checkout")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "libburyn_amd.so")), reason="engine library not built")
def test_shim_compiles_and_links_against_reference_prototypes(tmp_path):
    protos = _reference_prototypes()
    names = [n for n, _ in protos]
    assert len(names) == 10, names
    decls = "#include <cstdint>\n#include \"buryn_amd.h\"\n" + "\n".join(p for _, p in protos) + "\n"
    (tmp_path / "decls.h").write_text(decls)
    (tmp_path / "shim.cpp").write_text("#include \"decls.h\"\n" + _shim_source())
    refs = ",\n".join(f"    reinterpret_cast<void*>(&{n})" for n in names)
    (tmp_path / "main.cpp").write_text("#include \"decls.h\"\nint main() {\n  void* launchers[] = {\n" + refs +
                                       "};\n  return launchers[0] == nullptr;\n}\n")
    inc = ["-I", os.path.join(REPO, "include"), "-I", str(tmp_path)]
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", *inc, "-c", str(tmp_path / "shim.cpp"), "-o",
                    str(tmp_path / "shim.o")], check=True, capture_output=True, text=True)
    r = subprocess.run(["g++", "-std=c++17", *inc, str(tmp_path / "main.cpp"), str(tmp_path / "shim.o"), "-L", LIB,
                        "-lburyn_amd", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(tmp_path / "ref_caller")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-C", str(tmp_path / "shim.o")], check=True, capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf" T {n}\(", nm), f"{n} not defined by the shim"


@pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference checkout")
def test_shim_signature_mismatch_fails_to_link(tmp_path):
    # The check itself: a shim whose evolve launcher takes a pointer instead of the view by value
    # leaves the reference's declaration undefined.
    protos = _reference_prototypes()
    decls = "#include <cstdint>\n#include \"buryn_amd.h\"\n" + "\n".join(p for _, p in protos) + "\n"
    (tmp_path / "decls.h").write_text(decls)
    bad = _shim_source().replace("void kernel_neuron_group_evolve_launch(NeuronGroupView v) {",
                                 "void kernel_neuron_group_evolve_launch(NeuronGroupView* pv) {\n    NeuronGroupView v = *pv;")
    assert bad != _shim_source()
    (tmp_path / "shim.cpp").write_text("#include \"decls.h\"\n" + bad)
    (tmp_path / "main.cpp").write_text("#include \"decls.h\"\nint main() { return reinterpret_cast<void*>("
                                       "&kernel_neuron_group_evolve_launch) == nullptr; }\n")
    inc = ["-I", os.path.join(REPO, "include"), "-I", str(tmp_path)]
    subprocess.run(["g++", "-std=c++17", *inc, "-c", str(tmp_path / "shim.cpp"), "-o", str(tmp_path / "shim.o")],
                   check=True, capture_output=True)
    r = subprocess.run(["g++", "-std=c++17", *inc, str(tmp_path / "main.cpp"), str(tmp_path / "shim.o"), "-o",
                        str(tmp_path / "x")], capture_output=True, text=True)
    assert r.returncode != 0 and "kernel_neuron_group_evolve_launch" in r.stderr
