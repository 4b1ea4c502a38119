"""Writes the ChBrClF golden fixture used by the parity tests.

Run in a checkout that has the reference tree mounted (it reads data only:
the numeric initialisers of ``tests/chbrclf.hpp`` and the numbers of
``tests/mo_reference.txt``), e.g.::

    python tests/golden/extract_chbrclf.py /root/reference

The output ``tests/golden/chbrclf.npz`` holds, as plain arrays (no pickles):

* the ChBrClF system the reference tests load (``tests/chbrclf.hpp``):
  nucleus charges/coordinates, the 14 AO-basis arrays, the MO coefficients
  ``[mo_num][ao_num]`` and the electron coordinates ``[walk][elec][3]``;
* ``mo_reference``: the full ``mo_vgl[136][5][224]`` that
  ``tests/test_qmckl_mo.cpp:382-405`` checks at 1e-14 absolute.

The 9 scalar known answers of ``tests/test_qmckl_ao.cpp:440-447`` and
``tests/test_qmckl_mo.cpp:250-259`` are short enough to sit in the tests
themselves.
"""
import hashlib
import os
import re
import sys

import numpy as np

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _array(text, name, dtype):
    m = re.search(r"\b" + re.escape(name) + r"\s*\[[^=]*=\s*\{(.*?)\}\s*;", text, re.S)
    if m is None:
        raise KeyError(name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return np.array([float(x) for x in _NUM.findall(body)]).astype(dtype)


def _define(text, name):
    m = re.search(r"#define\s+" + re.escape(name) + r"\s+\(?\s*(?:\(int64_t\))?\s*(\d+)", text)
    if m is None:
        raise KeyError(name)
    return int(m.group(1))


def main(ref_root):
    hpp_path = os.path.join(ref_root, "tests", "chbrclf.hpp")
    txt_path = os.path.join(ref_root, "tests", "mo_reference.txt")
    with open(hpp_path) as f:
        text = f.read()
    nucl_num = _define(text, "chbrclf_nucl_num")
    shell_num = _define(text, "chbrclf_shell_num")
    prim_num = _define(text, "chbrclf_prim_num")
    ao_num = _define(text, "chbrclf_ao_num")
    mo_num = _define(text, "chbrclf_mo_num")
    elec_up = _define(text, "chbrclf_elec_up_num")
    elec_dn = _define(text, "chbrclf_elec_dn_num")
    walk_num = _define(text, "chbrclf_walk_num")
    elec_num = elec_up + elec_dn

    out = dict(
        nucl_num=np.int64(nucl_num), shell_num=np.int64(shell_num),
        prim_num=np.int64(prim_num), ao_num=np.int64(ao_num), mo_num=np.int64(mo_num),
        elec_up_num=np.int64(elec_up), elec_dn_num=np.int64(elec_dn),
        walk_num=np.int64(walk_num),
        charge=_array(text, "chbrclf_charge", np.float64),
        nucl_coord_T=_array(text, "chbrclf_nucl_coord", np.float64).reshape(3, nucl_num),
        nucleus_index=_array(text, "chbrclf_basis_nucleus_index", np.int64),
        nucleus_shell_num=_array(text, "chbrclf_basis_nucleus_shell_num", np.int64),
        shell_ang_mom=_array(text, "chbrclf_basis_shell_ang_mom", np.int32),
        shell_prim_num=_array(text, "chbrclf_basis_shell_prim_num", np.int64),
        shell_prim_index=_array(text, "chbrclf_basis_shell_prim_index", np.int64),
        shell_factor=_array(text, "chbrclf_basis_shell_factor", np.float64),
        ao_factor=_array(text, "chbrclf_basis_ao_factor", np.float64),
        exponent=_array(text, "chbrclf_basis_exponent", np.float64),
        coefficient=_array(text, "chbrclf_basis_coefficient", np.float64),
        prim_factor=_array(text, "chbrclf_basis_prim_factor", np.float64),
        mo_coef=_array(text, "chbrclf_mo_coef", np.float64).reshape(mo_num, ao_num),
        elec_coord=_array(text, "chbrclf_elec_coord", np.float64).reshape(walk_num, elec_num, 3),
    )
    sizes = dict(charge=nucl_num, nucleus_index=nucl_num, nucleus_shell_num=nucl_num,
                 shell_ang_mom=shell_num, shell_prim_num=shell_num,
                 shell_prim_index=shell_num, shell_factor=shell_num, ao_factor=ao_num,
                 exponent=prim_num, coefficient=prim_num, prim_factor=prim_num)
    for k, n in sizes.items():
        assert out[k].shape == (n,), (k, out[k].shape, n)

    with open(txt_path) as f:
        ref = np.array([float(x) for x in f.read().split()], dtype=np.float64)
    point_num = walk_num * elec_num
    assert ref.size == point_num * 5 * mo_num, ref.size
    out["mo_reference"] = ref.reshape(point_num, 5, mo_num)

    sha = {}
    for p in (hpp_path, txt_path):
        with open(p, "rb") as f:
            sha[os.path.basename(p)] = hashlib.sha256(f.read()).hexdigest()
    out["source_sha256"] = np.array([f"{k}:{v}" for k, v in sorted(sha.items())])

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chbrclf.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, {k: getattr(v, "shape", ()) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
