"""Writes the golden fixtures under tests/golden/ (run once, in a container
that holds the reference at /root/reference or $LYWELL_REFERENCE).

1. Data files the reference ships with its own known-answer runs
   (Release_Intel64/), converted to compact .npz:
   - winterp_sphere.npz: the 512x242 interpolation matrix dumped by
     test_interp (modMLFMA.f90:1637-1779) for sphere.stl at lambda = 0.95 m
     (only its nonzeros), the two sample grids, the field samples before and
     after Interpolar/Anterpolar, and the sphere.stl surface itself;
   - ilu0_saad.npz: the 10x10 COO matrix {i,j,s,si}.dat and the ILU0 factor
     testSAAD wrote (main.f90:719-826, ilu0_{ao_real,ao_imag,ir,jc}).
2. sphere.stl itself (gzip) and the arrays the reference's STL reader
   returns for it (stl_sphere_ref.npz).
3. Golden vectors from the flang build of the reference (oracle/_ref,
   oracle/build_ref.sh) on the 1,920-unknown sphere (radius 0.5 m, icosphere
   subdivision 3, lambda = 1 m, ladomax = 0.2 -> 4 levels):
   - ref_sphere1920.npz: the mesh fed to the reference, its RWG arrays,
     MatxVec_MLFMA of the deterministic vector and of the plane wave, a sample
     of the near field and of the ILU0 factor, lusol0C, and CGSold + sigmaRCS
     with and without ILU0 (matvec counts included).
4. dat_writers.npz: the text the reference's .dat writers produce (flang
   build, oracle/ref_dat.f90) for values that probe the F16.8 and I10 fields,
   and samples of .dat files the reference's own runs wrote.
Run with no argument for everything, or name parts: files stl build dat.
"""
import os
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("LYWELL_REFERENCE", "/root/reference")
REL = os.path.join(REF, "Release_Intel64")
sys.path.insert(0, ROOT)


def load_cols(name):
    return np.loadtxt(os.path.join(REL, name))


def read_stl(path):
    pts, tris, nrm = [], [], []
    index = {}
    cur = []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            if w[0] == "facet" and w[1] == "normal":
                nrm.append([float(v) for v in w[2:5]])
            elif w[0] == "vertex":
                key = tuple(float(v) for v in w[1:4])
                if key not in index:
                    index[key] = len(pts) + 1
                    pts.append(key)
                cur.append(index[key])
                if len(cur) == 3:
                    tris.append(cur)
                    cur = []
    return np.array(pts), np.array(tris, np.int32), np.array(nrm)


def fixtures_from_reference_files():
    W = load_cols("_Winterp.dat")
    r, c = np.nonzero(W)
    comp = {}
    for part in ("orig", "inter", "anter", "met1", "met2"):
        for pol in ("t", "p"):
            re = np.atleast_2d(load_cols(f"_{part}_{pol}_R.dat"))
            im = np.atleast_2d(load_cols(f"_{part}_{pol}_I.dat"))
            comp[f"{part}_{pol}"] = re + 1j * im
    p, t, n = read_stl(os.path.join(REF, "sphere.stl"))
    np.savez_compressed(os.path.join(HERE, "winterp_sphere.npz"), shape=np.array(W.shape), rows=r.astype(np.int32),
                        cols=c.astype(np.int32), vals=W[r, c], thetas_n=load_cols("_thetas_n.dat"),
                        phis_n=load_cols("_phis_n.dat"), thetas_n1=load_cols("_thetas_n1.dat"),
                        phis_n1=load_cols("_phis_n1.dat"), stl_p=p, stl_t=t, stl_n=n, lam=0.95, ladomax=0.5,
                        **comp)

    def col(name):
        v = np.loadtxt(os.path.join(REL, name))
        return v[1:]  # first line holds the count

    ao = np.loadtxt(os.path.join(REL, "ilu0_ao_real")) + 1j * np.loadtxt(os.path.join(REL, "ilu0_ao_imag"))
    np.savez_compressed(os.path.join(HERE, "ilu0_saad.npz"), i=col("i.dat").astype(np.int32),
                        j=col("j.dat").astype(np.int32), s=col("s.dat"), si=col("si.dat"), ao=ao,
                        ir=np.loadtxt(os.path.join(REL, "ilu0_ir")).astype(np.int32),
                        jc=np.loadtxt(os.path.join(REL, "ilu0_jc")).astype(np.int32))


def run_ref_driver(p, t, nrm, lam, lad, errmax, theta, phi, d, pol, ctte, use_prec):
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "in.bin"), "wb") as f:
            np.array([p.size // 3, t.size // 3], np.int32).tofile(f)
            p.tofile(f)
            t.tofile(f)
            nrm.tofile(f)
            np.array([lam, lad, errmax]).tofile(f)
            np.array([len(theta), use_prec], np.int32).tofile(f)
            np.asarray(d, np.float64).tofile(f)
            np.asarray(pol, np.complex128).tofile(f)
            np.array([ctte], np.complex128).tofile(f)
            theta.tofile(f)
            np.array([phi]).tofile(f)
        subprocess.run([drv, "in.bin", "out.bin"], cwd=td, check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(td, "out.bin"), "rb") as f:
            def rd(dt, n):
                return np.fromfile(f, dt, n)
            ne, nnz = (int(v) for v in rd(np.int32, 2))
            o = dict(ne=ne, nnz=nnz)
            o["e_t"] = rd(np.int32, 2 * ne)
            o["e_po"] = rd(np.int32, 2 * ne)
            o["e_long"] = rd(np.float64, ne)
            o["y1"] = rd(np.complex128, ne)
            o["y2"] = rd(np.complex128, ne)
            o["ia"] = rd(np.int32, ne + 1)
            o["ja"] = rd(np.int32, nnz)
            o["z"] = rd(np.complex64, nnz)
            o["ierr"] = int(rd(np.int32, 1)[0])
            o["alu"] = rd(np.complex64, nnz + 1)
            o["jlu"] = rd(np.int32, nnz + 1)
            o["ju"] = rd(np.int32, ne)
            o["lus"] = rd(np.complex128, ne)
            o["b"] = rd(np.complex128, ne)
            o["nmv"] = int(rd(np.int32, 1)[0])
            o["x"] = rd(np.complex128, ne)
            o["sigma"] = rd(np.float64, len(theta))
    return o


def golden_from_reference_build():
    import lywell_amd as lw
    mesh = lw.Mesh.icosphere(3, 0.5)
    lam, lad, errmax = 1.0, 0.2, 1e-4
    theta = np.linspace(0.0, np.pi, 181)
    d, pol, ctte = [0.0, 0.0, -1.0], [1.0, 0.0, 0.0], 1.0
    runs = {}
    for use_prec in (0, 1):
        print(f"reference run, ILU0 = {use_prec}", flush=True)
        runs[use_prec] = run_ref_driver(mesh.p_coord, mesh.t_p, mesh.t_normal, lam, lad, errmax, theta, 0.0, d, pol,
                                        ctte, use_prec)
    o = runs[1]
    step = 97
    np.savez_compressed(
        os.path.join(HERE, "ref_sphere1920.npz"), p=mesh.p_coord, t=mesh.t_p, nrm=mesh.t_normal, lam=lam, ladomax=lad,
        errmax=errmax, theta=theta, phi=0.0, d=np.array(d), pol=np.array(pol, np.complex128), ctte=ctte,
        e_t=o["e_t"], e_po=o["e_po"], e_long=o["e_long"], y1=o["y1"], y2=o["y2"], b=o["b"], nnz=o["nnz"],
        ia=o["ia"], ja_crc=zlib.crc32(o["ja"].tobytes()), z_step=step, z_sample=o["z"][::step], ierr=o["ierr"],
        jlu_crc=zlib.crc32(o["jlu"].tobytes()), ju=o["ju"], alu_sample=o["alu"][::step], lus=o["lus"],
        nmv_prec=runs[1]["nmv"], x_prec=runs[1]["x"], sigma_prec=runs[1]["sigma"], nmv_noprec=runs[0]["nmv"],
        x_noprec=runs[0]["x"], sigma_noprec=runs[0]["sigma"])


def run_ref_stl(path):
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_stl"), path, os.path.join(td, "o.bin")],
                       check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(td, "o.bin"), "rb") as f:
            np_, nt = (int(v) for v in np.fromfile(f, np.int32, 2))
            return np.fromfile(f, np.float64, 3 * np_), np.fromfile(f, np.int32, 3 * nt), np.fromfile(f, np.float64, 3 * nt)


def stl_fixture():
    """sphere.stl as shipped (gzip) and what the reference's reader returns for
    it (obtener_p_t: node order after limpiarnodos, triangles, normals)."""
    import gzip
    src = os.path.join(REF, "sphere.stl")
    with open(src, "rb") as f, gzip.open(os.path.join(HERE, "sphere.stl.gz"), "wb") as g:
        g.write(f.read())
    p, t, n = run_ref_stl(src)
    np.savez_compressed(os.path.join(HERE, "stl_sphere_ref.npz"), p=p, t=t, n=n)


def dat_writer_inputs():
    """Values that probe F16.8 and I10: signed zeros, negative values that round
    to zero, exact decimal ties, the field-width limits, NaN and infinities,
    then random values over 16 decades."""
    rng = np.random.default_rng(0)
    special = [0.0, -0.0, 1e-10, -1e-10, -4.9e-9, -5.1e-9, 4.9999999e-9, 5e-9, 1 / 512, -1 / 512, 3 / 1024, 0.5, -0.5,
               1.0, -1.0, 9999999.0, 9999999.999999994, 9999999.999999996, -999999.99999999, -9999999.0, 12345678.9,
               1e300, -1e300, 123.456789125, 2.0 ** -27, 0.125000005, np.pi, -np.e, np.nan, np.inf, -np.inf, 2.5e-9]
    vals = np.concatenate([special, rng.standard_normal(200) * 10.0 ** rng.integers(-9, 7, 200),
                           rng.uniform(-1, 1, 60) * 1e-8])
    n = len(vals) - len(vals) % 6
    vals = vals[:n].astype(np.float64)
    iv = np.concatenate([[0, -1, 1, 2 ** 31 - 1, -2 ** 31, 999999999, -999999999, 1234567890, -123456789],
                         rng.integers(-10 ** 9, 10 ** 9, n - 9)]).astype(np.int32)
    return vals, iv


def dat_writer_fixture():
    """What the reference's guardarVectorC / guardarMatrizR / guardarVectorR /
    guardarVectorI (modGlobalMetod.f90:186-306, flang build) write for
    dat_writer_inputs(), and samples of the .dat files the reference's own
    (ifort) runs wrote: a current (conkind8_IMLFMA.dat, guardarVectorC) and
    monostatic tables (guardarMatrizR)."""
    vals, iv = dat_writer_inputs()
    texts = {}
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "in.bin"), "wb") as f:
            np.array([vals.size], np.int32).tofile(f)
            vals.tofile(f)
            iv.tofile(f)
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_dat"), "in.bin", "out"], cwd=td, check=True)
        for k in ("vc", "mr", "vr", "vi"):
            with open(os.path.join(td, f"out_{k}.dat")) as f:
                texts["flang_" + k] = np.array(f.read())
    for name, key, lines in (("conkind8_IMLFMA.dat", "ifort_current", 3000),
                             ("pruebaVH_RCSMONOVV.dat", "ifort_mono_vv", None),
                             ("last_rcs_mono_incremental.dat", "ifort_mono_incremental", None)):
        with open(os.path.join(REL, name)) as f:
            txt = f.read()
        if lines:
            txt = "".join(txt.splitlines(keepends=True)[:lines])
        texts[key] = np.array(txt)
    np.savez_compressed(os.path.join(HERE, "dat_writers.npz"), vals=vals, iv=iv, **texts)


if __name__ == "__main__":
    only = sys.argv[1:]
    if not only or "files" in only:
        fixtures_from_reference_files()
    if not only or "stl" in only:
        stl_fixture()
    if not only or "build" in only:
        golden_from_reference_build()
    if not only or "dat" in only:
        dat_writer_fixture()
    print("fixtures written to", HERE)
