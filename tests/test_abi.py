"""The C boundary: ABI v13 record layout and the exported entry points."""

import re
import subprocess

import numpy as np
import pytest

from conftest import REPO

# Record sizes asserted by the reference header (include/numi/cloth_bag_gpu.h:192-209).
REFERENCE_SIZES = {
    "NumiClothBagGPUConfig": 224, "NumiClothBagGPUParticle": 48, "NumiClothBagGPUDistance": 32,
    "NumiClothBagGPUGrip": 48, "NumiClothBagGPUKnot": 48, "NumiClothBagGPUBend": 32,
    "NumiClothBagGPUFruit": 96, "NumiClothBagGPUFruitPair": 32, "NumiClothBagGPUYarnContact": 112,
    "NumiClothBagGPUSelfPair": 8, "NumiClothBagGPUSelfStatus": 16, "NumiClothBagGPUSelfImpulse": 48,
    "NumiClothBagGPUFrictionStatus": 32, "NumiClothBagGPUAerodynamicsStatus": 16,
    "NumiClothBagGPUReleaseStatus": 16, "NumiClothBagGPUBatch": 16,
}


def _compile_probe(tmp_path, body: str) -> str:
    source = tmp_path / "probe.cpp"
    source.write_text("#include <cstddef>\n#include <cstdio>\n#include \"numi_cloth_hip.h\"\n"
                      "#include \"numi_cloth_scene.h\"\nint main() {\n" + body + "\nreturn 0; }\n")
    binary = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", f"-I{REPO / 'include'}", str(source), "-o", str(binary)], check=True)
    return subprocess.run([str(binary)], check=True, capture_output=True, text=True).stdout


def test_record_sizes_match_reference(tmp_path):
    body = "\n".join(f'std::printf("{name} %zu\\n", sizeof({name}));' for name in REFERENCE_SIZES)
    sizes = dict(line.split() for line in _compile_probe(tmp_path, body).splitlines())
    assert {k: int(v) for k, v in sizes.items()} == REFERENCE_SIZES


def test_python_dtypes_match_c_offsets(tmp_path, cloth):
    records = {
        "NumiClothBagGPUConfig": cloth.CONFIG, "NumiClothBagGPUParticle": cloth.PARTICLE,
        "NumiClothBagGPUDistance": cloth.DISTANCE, "NumiClothBagGPUGrip": cloth.GRIP,
        "NumiClothBagGPUKnot": cloth.KNOT, "NumiClothBagGPUBend": cloth.BEND, "NumiClothBagGPUFruit": cloth.FRUIT,
        "NumiClothBagGPUFruitPair": cloth.FRUIT_PAIR, "NumiClothBagGPUYarnContact": cloth.YARN_CONTACT,
        "NumiClothBagGPUBatch": cloth.BATCH,
    }
    lines = []
    for struct, dtype in records.items():
        for field in dtype.names:
            lines.append(f'std::printf("{struct}.{field} %zu\\n", offsetof({struct}, {field}));')
    printed = dict(line.split() for line in _compile_probe(tmp_path, "\n".join(lines)).splitlines())
    for struct, dtype in records.items():
        for field in dtype.names:
            assert int(printed[f"{struct}.{field}"]) == dtype.fields[field][1], (struct, field)


def test_python_structs_match_c_layout(tmp_path, cloth):
    import ctypes
    structs = {"NumiClothHipSceneDesc": cloth.cloth.SceneDesc, "NumiClothHipState": cloth.cloth.State,
               "NumiClothHipDiagnostics": cloth.cloth.Diagnostics,
               "NumiClothHipSceneExtension": cloth.cloth.SceneExtension,
               "NumiClothHipSheetSpec": cloth.cloth.SheetSpec, "NumiClothHipBagSpec": cloth.cloth.BagSpec,
               "NumiClothMaterialArtifact": cloth.cloth.MaterialArtifact,
               "NumiClothGripPose": cloth.cloth.GripPose,
               "NumiClothGripTrajectoryInfo": cloth.cloth.GripTrajectoryInfo}
    lines = []
    for name, struct in structs.items():
        lines.append(f'std::printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in struct._fields_:
            lines.append(f'std::printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    printed = dict(line.split() for line in _compile_probe(tmp_path, "\n".join(lines)).splitlines())
    for name, struct in structs.items():
        assert int(printed[name]) == ctypes.sizeof(struct), name
        for field, _ in struct._fields_:
            assert int(printed[f"{name}.{field}"]) == getattr(struct, field).offset, (name, field)


def _declared_functions():
    names = set()
    for header in (REPO / "include").glob("*.h"):
        text = header.read_text()
        names.update(re.findall(r"^[A-Za-z_][\w \*]*?\b(numi_cloth_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_entry_point(cloth):
    declared = _declared_functions()
    assert len(declared) >= 20
    exported = subprocess.run(["nm", "-D", "--defined-only", str(cloth.cloth.LIBRARY_PATH)], check=True,
                              capture_output=True, text=True).stdout
    symbols = {line.split()[-1] for line in exported.splitlines() if line.strip()}
    missing = [name for name in declared if name not in symbols]
    assert not missing, missing


def test_abi_version(cloth):
    assert cloth.library().numi_cloth_hip_abi_version() == 13


def test_null_arguments_are_rejected_without_a_gpu(cloth):
    lib = cloth.library()
    assert lib.numi_cloth_hip_create(None, None) == 1
    assert b"null" in lib.numi_cloth_hip_last_error()
    assert lib.numi_cloth_hip_step(None, None, 1, 1, 1) == 1
    assert lib.numi_cloth_hip_destroy(None) == 0
