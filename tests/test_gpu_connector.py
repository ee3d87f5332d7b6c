"""The reference's Python/MATLAB connector entry points (py_connector.c:63,
:146) on the GPU path: every exact algorithm ID (4-8) on files of each
element type, against the exact oracle."""
import numpy as np
import pytest

import oracle as O
import permanent_amd as P

pytestmark = pytest.mark.gpu


def write_matrix(path, a, type_name):
    nz = [(i, j, a[i, j]) for i in range(a.shape[0]) for j in range(a.shape[1]) if a[i, j] != 0]
    fmt = {"int": "{:d}", "float": "{:.9g}", "double": "{:.17g}"}[type_name]
    lines = [f"{a.shape[0]} {len(nz)} {type_name}"]
    lines += [f"{i} {j} " + fmt.format(int(v) if type_name == "int" else float(v))
              for i, j, v in nz]
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("type_name", ["int", "float", "double"])
@pytest.mark.parametrize("algo", [4, 5, 6, 7, 8])
def test_read_calculate_return(tmp_path, type_name, algo):
    a = P.synthetic_matrix(14, 0.6, seed=61)
    if type_name == "int":
        a = np.round(a * 3).astype(np.int32)
    elif type_name == "float":
        a = a.astype(np.float32)
    f = tmp_path / f"m_{type_name}.txt"
    write_matrix(f, a, type_name)
    parsed = O.read_matrix(str(f))[0]  # the reader restatement (util.h:343-358)
    ex = float(O.exact_permanent(np.asarray(parsed, dtype=np.float64)))
    v = P.read_calculate_return(str(f), algo)
    assert v == pytest.approx(ex, rel=1e-12), (type_name, algo, v, ex)
    if type_name == "int" and abs(ex) < 2 ** 53:
        assert v == ex


def test_matlab_calculate_return():
    a = P.synthetic_matrix(12, 0.5, seed=62, binary=True).astype(np.int32)
    ex = float(O.exact_permanent(a))
    for algo in (4, 5, 6, 7, 8):
        assert P.matlab_calculate_return(a, algo) == ex


def _lib_connect():
    """libConnect.so bound the way the reference's callers bind it:
    superpython.py loads it with cdll.LoadLibrary and sets the argtypes of
    read_calculate_return (superpython.py:6-26); supermaTlab.m calls
    matlab_calculate_return through matlab_calculate_return.h."""
    import ctypes
    import os

    lib = ctypes.cdll.LoadLibrary(os.path.join(os.path.dirname(P.LIB_PATH), "libConnect.so"))
    lib.connect()
    lib.read_calculate_return.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.read_calculate_return.restype = ctypes.c_double
    lib.matlab_calculate_return.argtypes = [ctypes.POINTER(ctypes.c_int)] + [ctypes.c_int] * 7
    lib.matlab_calculate_return.restype = ctypes.c_double
    return lib


@pytest.mark.parametrize("algo", [4, 5, 6, 7, 8])
def test_lib_connect_under_the_reference_names(tmp_path, algo):
    """The reference's callers switch by changing only the library path:
    libConnect.so's connect / read_calculate_return / matlab_calculate_return
    give the exact permanent for every exact algorithm ID."""
    import ctypes

    lib = _lib_connect()
    a = P.synthetic_matrix(15, 0.55, seed=63)
    f = tmp_path / "m.txt"
    write_matrix(f, a, "double")
    ex = float(O.exact_permanent(a))
    v = lib.read_calculate_return(str(f).encode("utf-8"), algo, 8, 10, 4, 5)
    assert v == pytest.approx(ex, rel=1e-12), (algo, v, ex)
    b = P.synthetic_matrix(11, 0.5, seed=64, binary=True).astype(np.int32)
    bp = np.ascontiguousarray(b)
    w = lib.matlab_calculate_return(bp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), algo,
                                    2, 2, 2, 2, 11, int((b != 0).sum()))
    assert w == float(O.exact_permanent(b))
