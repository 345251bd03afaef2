"""The C-ABI library loads, exports every symbol of include/*.h, and its struct layout matches the bindings."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt, flags=re.M):
            name = m.group(1)
            if name not in ("static_assert",):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported(lib):
    from fltk_sph_ice_amd._abi import CXX_SYMBOLS, LIB_PATH, SIGNATURES

    declared = _declared()
    assert len(declared) > 50
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    unbound = sorted(declared - set(SIGNATURES) - set(CXX_SYMBOLS))
    assert not unbound, f"declared but not bound in _abi: {unbound}"
    # the C++-linkage launch functions keep the reference's C++ signatures (mangled names)
    import subprocess

    nm = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True).stdout
    for name, (mangled, _) in CXX_SYMBOLS.items():
        assert mangled in nm, name
    demangled = subprocess.run(["c++filt"], input="\n".join(m for m, _ in CXX_SYMBOLS.values()), capture_output=True,
                               text=True).stdout
    assert "LaunchShapeMatchingGPU(int, float*, float*, float*, float*, float*, float*, float*, int*, int*, float)" in demangled


def test_struct_layout_matches_bindings(lib):
    from fltk_sph_ice_amd._abi import rxParticleCell, rxSimParams

    assert lib.rxAbiLayout(0) == C.sizeof(rxSimParams)
    assert lib.rxAbiLayout(1) == C.sizeof(rxParticleCell)
    assert lib.rxAbiLayout(2) == rxSimParams.EffectiveRadius.offset
    assert lib.rxAbiLayout(3) == rxSimParams.AP.offset
    assert lib.rxAbiLayout(4) == rxSimParams.BoxNum.offset
    assert lib.rxAbiLayout(5) == rxSimParams.SphereFlg.offset
    assert lib.rxAbiLayout(6) == rxParticleCell.uNumParticles.offset
    assert lib.rxAbiLayout(7) == rxParticleCell.uNumPolyHash.offset


def test_reference_struct_size():
    # rx_cu_common.cuh:142-237 with float3 = 12 bytes, all members 4-byte aligned:
    # 308 bytes of scalar/vector members up to AP_WQ, PolyNum + PolyVel[10] (124), BoxNum + BoxCen/BoxExt[10]
    # (244) + BoxRot/BoxInvRot[10] (720) + BoxFlg[10] (40), SphereNum + SphereCen[10] + SphereRad[10] +
    # SphereFlg[10] (204) = 1640 bytes, no padding.
    from fltk_sph_ice_amd._abi import rxSimParams

    assert C.sizeof(rxSimParams) == 1640


def test_out_of_scope_entry_points_report_an_error(lib):
    # meshing / turbulence / sub-particle / screen-space entry points link and report error 5 without work
    lib.CuClearLastError()
    lib.CuSubInit()
    assert lib.CuGetLastErrorCode() == 5
    assert b"CuSubInit" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
    lib.CuMCCreateMesh()
    assert lib.CuGetLastErrorCode() == 5
    lib.CuClearLastError()
    assert lib.CuCheckNumValidHashData() == 0
    lib.CuClearLastError()


def test_path_shape_matching_links_with_the_reference_signatures_and_reports_an_error(lib):
    # Surf_SM.h:34-52, Ice_SM.h:55-69 and IceObject.h:399-425 (C++ linkage): the stubs take the reference's argument lists and touch
    # no memory, so null pointers are safe here
    import subprocess

    from fltk_sph_ice_amd._abi import CXX_SYMBOLS

    demangled = subprocess.run(["c++filt"], input="\n".join(CXX_SYMBOLS[n][0] for n in ("LaunchShapeMatchingUsePathGPU",
                                                                                     "LauchUpdateSMFromPath")),
                               capture_output=True, text=True).stdout.splitlines()
    assert demangled[0] == ("LaunchShapeMatchingUsePathGPU(int, float*, float*, float*, float*, float*, float*, float*, float*, "
                            "int*, int*, float)")
    assert demangled[1] == ("LauchUpdateSMFromPath(int, float*, float*, float*, float*, float*, float*, float*, float*, int*, "
                            "int*, short*, short*, float*, float*, int*, int*, int, int, float)")
    lib.CuClearLastError()
    lib.LaunchShapeMatchingUsePathGPU(4, *([None] * 10), C.c_float(0.01))
    assert lib.CuGetLastErrorCode() == 5
    assert b"LaunchShapeMatchingUsePathGPU" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
    lib.LauchUpdateSMFromPath(4, *([None] * 17), 8, 2, C.c_float(0.01))
    assert lib.CuGetLastErrorCode() == 5
    assert b"LauchUpdateSMFromPath" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
    lib.LaunchUpdatePrefixSumGPU(4, 1, 1, 1, 1, *([None] * 11))
    assert lib.CuGetLastErrorCode() == 5
    assert b"LaunchUpdatePrefixSumGPU" in lib.CuGetLastErrorString()
    lib.CuClearLastError()
    assert subprocess.run(["c++filt", CXX_SYMBOLS["LaunchUpdatePrefixSumGPU"][0]], capture_output=True, text=True).stdout.strip() == (
        "LaunchUpdatePrefixSumGPU(int, int, int, int, int, int*, short*, float*, float*, short*, float*, float*, unsigned int*, "
        "unsigned int*, float const*, float const*)")
