"""Writes the staged-input HDF5 fixtures for tests/test_staged_h5.py with the real HDF5 library (h5py).

Run with an interpreter that has h5py (the trot_amd interpreter does not):

    PYTHONPATH=/opt/conda/lib/python3.9/site-packages /opt/conda/bin/python3.9 tests/golden/make_staged_h5.py

The files follow the reference's cache layout (trot/staging.py:1353-1376, _dump_h5 and _dump_frozen):
root attr meta_json; /ham datasets h0, h1, chol, nelec and attrs norb, chol_cut, source_kind, basis, frozen
(int attr, or an int64 dataset for an array); /trial attrs kind, source_kind, frozen; /trial/data/<arrays>.
Two file-format generations are written (h5py's default "earliest": superblock 0, version-1 object
headers, symbol-table groups; and "latest": superblock 3, version-2 object headers, link messages),
plus one file with a chunked, compressed dataset that the reader must refuse.  The expected values go to
staged_expected.npz.
"""

import json
from pathlib import Path

import h5py
import numpy as np

HERE = Path(__file__).resolve().parent


def system(seed, N, G):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((N, N))
    h1 = 0.5 * (a + a.T)
    b = rng.standard_normal((G, N, N))
    chol = 0.5 * (b + np.swapaxes(b, 1, 2))
    mo = np.linalg.qr(rng.standard_normal((N, N)))[0]
    mob = np.linalg.qr(rng.standard_normal((N, N)))[0]
    return float(rng.standard_normal()), h1, chol, mo, mob


def dump(path, libver, h0, h1, chol, nelec, frozen, kind, data, meta, chunked=False):
    with h5py.File(path, "w", libver=libver) as f:
        f.attrs["meta_json"] = json.dumps(meta)
        g = f.create_group("ham")
        g.create_dataset("h0", data=np.array(h0))
        g.create_dataset("h1", data=h1)
        if chunked:
            g.create_dataset("chol", data=chol, chunks=(1,) + chol.shape[1:], compression="gzip")
        else:
            g.create_dataset("chol", data=chol)
        g.create_dataset("nelec", data=np.array(nelec, dtype=np.int64))
        g.attrs["norb"] = h1.shape[0]
        g.attrs["chol_cut"] = 1e-5
        if isinstance(frozen, np.ndarray):
            g.create_dataset("frozen", data=np.asarray(frozen, dtype=np.int64))
        else:
            g.attrs["frozen"] = int(frozen)
        g.attrs["source_kind"] = "mf"
        g.attrs["basis"] = "restricted"
        t = f.create_group("trial")
        t.attrs["kind"] = kind
        if isinstance(frozen, np.ndarray):
            t.create_dataset("frozen", data=np.asarray(frozen, dtype=np.int64))
        else:
            t.attrs["frozen"] = int(frozen)
        t.attrs["source_kind"] = "mf"
        d = t.create_group("data")
        for k, v in data.items():
            d.create_dataset(k, data=np.asarray(v))


def main():
    exp = {}
    h0, h1, chol, mo, mob = system(0, 6, 7)
    dump(HERE / "staged_rhf_earliest.h5", "earliest", h0, h1, chol, (3, 3), 0, "rhf", {"mo": mo},
         {"frozen": 0, "basis": "sto-6g", "note": "rhf fixture"})
    exp.update(rhf_h0=h0, rhf_h1=h1, rhf_chol=chol, rhf_mo=mo)
    h0, h1, chol, mo, mob = system(1, 5, 9)
    frozen = np.array([0, 7], dtype=np.int64)
    dump(HERE / "staged_uhf_latest.h5", "latest", h0, h1, chol, (3, 2), frozen, "uhf", {"mo_a": mo, "mo_b": mob},
         {"frozen": [0, 7], "note": "uhf fixture"})
    exp.update(uhf_h0=h0, uhf_h1=h1, uhf_chol=chol, uhf_mo_a=mo, uhf_mo_b=mob, uhf_frozen=frozen)
    dump(HERE / "staged_chunked.h5", "earliest", h0, h1, chol, (3, 2), 0, "rhf", {"mo": mo}, {}, chunked=True)
    np.savez(HERE / "staged_expected.npz", **exp)


if __name__ == "__main__":
    main()
