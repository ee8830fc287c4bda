"""Runtime compilation (hipRTC) through the C ABI: compilation needs no GPU, so these run on the
CPU. Options reach the compiler, external headers are found, failures carry the compiler log,
lowered names are recorded for kernels, template instantiations and globals."""
import os

import pytest

from superterrainplus_amd import capi

import rtc_matrix as M

SCRIPT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "superterrainplus_amd", "rtc",
                      "STPMultiHeightGenerator.rtc")


def test_missing_option_fails_with_log():
    with pytest.raises(capi.STPError) as e:
        capi.RtcProgram("MatrixOperator", M.SOURCE, ["-std=c++17"], M.NAMES)
    assert e.value.code == 5 and "STP_TEST_ENABLE" in str(e.value)


def test_compiles_and_lowers_names():
    p = capi.RtcProgram("MatrixOperator", M.SOURCE, ["-std=c++17", "-DSTP_TEST_ENABLE"], M.NAMES)
    # a global variable keeps its name; functions and template instantiations are mangled
    assert p.lowered("MatrixDimension") == "MatrixDimension"
    lowered = [p.lowered(n) for n in M.NAMES[1:]]
    assert all(n.startswith("_Z") for n in lowered) and len(set(lowered)) == 3
    assert "scale" in lowered[2]
    assert isinstance(p.log(), str)
    with pytest.raises(capi.STPError):
        p.lowered("notDeclared")


def test_external_header_is_found():
    """The program includes the attached header and stops with '#error PASS' after printing its
    version through '#pragma message'."""
    with pytest.raises(capi.STPError) as e:
        capi.RtcProgram("MatrixOperator", M.SOURCE, ["-std=c++17", "-DSTP_TEST_ENABLE", "-DSTP_TEST_EXTERNAL_HEADER"],
                        M.NAMES, headers={M.HEADER_NAME: M.HEADER})
    assert e.value.code == 5 and "4.2.1" in str(e.value) and "PASS" in str(e.value)


def test_header_missing_from_database_fails():
    with pytest.raises(capi.STPError) as e:
        capi.RtcProgram("MatrixOperator", M.SOURCE, ["-DSTP_TEST_ENABLE", "-DSTP_TEST_EXTERNAL_HEADER"], M.NAMES)
    assert e.value.code == 5


def test_syntax_error_reports_line():
    with pytest.raises(capi.STPError) as e:
        capi.RtcProgram("Broken", "__global__ void k(float* p) { p[0] = ; }\n")
    assert e.value.code == 5 and "error" in str(e.value)


def test_builtin_generator_script_compiles_with_library_headers():
    names = ["generateMultiBiomeHeightmap", "Dimension", "HalfDimension", "Permutation", "BiomeCount", "BiomeTable"]
    p = capi.RtcProgram("STPMultiHeightGenerator", open(SCRIPT).read(), ["-std=c++17", "-DSTP_BIOME_CAPACITY=16"],
                        names, library_headers=True)
    assert p.lowered("generateMultiBiomeHeightmap").startswith("_Z")
    assert p.lowered("BiomeTable") == "BiomeTable"
    # without the library headers the script's includes are not found
    with pytest.raises(capi.STPError):
        capi.RtcProgram("STPMultiHeightGenerator", open(SCRIPT).read(), ["-std=c++17"], names)


def test_relocatable_program_has_bitcode_only():
    """-fgpu-rdc gives LLVM bitcode (linked later) and no code object; linking needs a device,
    so the link itself is covered by the GPU tests."""
    p = capi.RtcProgram("MatrixOperator", M.SOURCE, ["-std=c++17", "-DSTP_TEST_ENABLE", "-fgpu-rdc"], M.NAMES)
    assert p.lowered("scale").startswith("_Z")


def test_builtin_splatmap_script_compiles_with_library_headers():
    """The splatmap script (SuperDemo+/Script/STPSplatmapGenerator.cu's contract) compiles against the
    library's hipRTC-safe headers and lowers its kernel and globals."""
    path = os.path.join(os.path.dirname(SCRIPT), "STPSplatmapGenerator.rtc")
    names = ["generateTextureSplatmap", "Dimension", "HalfDimension", "Permutation", "SplatDatabase"]
    p = capi.RtcProgram("STPSplatmapGenerator", open(path).read(), ["-std=c++17"], names, library_headers=True)
    assert p.lowered("generateTextureSplatmap").startswith("_Z")
    assert p.lowered("SplatDatabase") == "SplatDatabase"
