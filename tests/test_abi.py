"""CPU checks of the drop-in boundary: libompl_amd.so loads without a GPU,
exports every function include/ompl_amd.h declares, and the product does not
reach into the oracle."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "ompl_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ompl_amd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ompl_amd_check_motions_dev", "ompl_amd_validate_states_dev", "ompl_amd_nn_knn_dev",
                 "ompl_amd_nn_radius_dev", "ompl_amd_sample_uniform_dev", "ompl_amd_check_motion"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from ompl_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding knows every declared function's signature
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def test_abi_version_matches_header_and_bindings():
    from ompl_amd import _lib

    text = (ROOT / "include" / "ompl_amd.h").read_text()
    header = int(re.search(r"#define OMPL_AMD_VERSION (\d+)", text).group(1))
    assert header == _lib.ABI_VERSION == _lib.load().ompl_amd_version()


def test_error_path_without_gpu():
    import ctypes as C

    from ompl_amd import _lib

    lib = _lib.load()
    # a null handle is rejected before any device work, with a message
    rc = lib.ompl_amd_env_destroy(None)
    assert rc == 0
    h = C.c_void_p()
    rc = lib.ompl_amd_nn_create(7, 3, C.byref(h))
    assert rc == -1 and b"metric" in lib.ompl_amd_last_error()
    # an environment's own launch shapes need an environment
    rc = lib.ompl_amd_env_set_tuning(None, None)
    assert rc == -1 and b"environment" in lib.ompl_amd_last_error()


def test_product_does_not_use_the_oracle():
    banned = ("import oracle", "from oracle", "liboracle", "oracle_py", "orc_")
    for p in list((ROOT / "ompl_amd").rglob("*.py")) + list((ROOT / "ompl_amd").rglob("*.hip")) + \
            list((ROOT / "ompl_amd").rglob("*.cpp")) + list((ROOT / "ompl_amd").rglob("*.hpp")):
        text = p.read_text()
        for b in banned:
            assert b not in text, (p, b)
