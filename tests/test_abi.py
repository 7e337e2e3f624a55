"""The C-ABI library loads and exports every symbol include/isph_abi.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "isph_abi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isph_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_entry_points():
    names = declared_symbols()
    for ref in ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis",
                "int_vel", "ppesolve", "comp_vel", "comp_pos", "cellshiftalt", "freesurf",
                "opt2_shift", "densityupdate", "timestep", "scalart_anisotropic",
                "scalarupdate"):
        assert "isph_" + ref in names


def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(str(pkg.library_path()))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_package_export_list_matches_header(pkg):
    assert set(pkg.EXPORTS) <= set(declared_symbols())
