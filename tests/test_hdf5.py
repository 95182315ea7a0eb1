"""HDF5 files in the OPS layout (libosbl_h5.so through opensbli_amd.hdf5, and
the iohdf5 / output_hdf5 surface of the problem scripts), on the CPU.

The checker is independent of this library: the image's separate Python 3.9
with h5py (run as a subprocess) reads what this library writes the way the
reference's plot scripts do (apps/Sod_shock_tube/plot.py:14-31), and writes
files the way the reference's helper does (helperfunctions.py:80-127) for this
library to read; h5dump compares the metadata of both writers.
"""
import json
import os
import subprocess

import numpy as np
import pytest

CONDA_PY = "/opt/conda/bin/python3.9"
H5DUMP = "/opt/conda/bin/h5dump"


def _h5lib():
    from opensbli_amd import build, hdf5
    if not os.path.exists(hdf5.H5_LIB_PATH):
        if not build.hdf5_available():
            pytest.skip("no HDF5 library in this image")
        build.build_h5()
    return hdf5


def _conda(script, *args):
    """Run a script under the image's h5py-enabled Python; its stdout is JSON."""
    if not os.path.exists(CONDA_PY):
        pytest.skip("no h5py-enabled Python in this image")
    env = {k: v for k, v in os.environ.items() if not k.startswith("PYTHON")}
    r = subprocess.run([CONDA_PY, "-c", script, *args], capture_output=True, text=True, env=env, cwd="/tmp")
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py is not importable by %s" % CONDA_PY)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


# reads a file as the reference's plot scripts do, and reports every attribute with its HDF5 type
READ_LIKE_PLOT = r"""
import sys, json, h5py, numpy
f = h5py.File(sys.argv[1], 'r')
out = {'groups': list(f.keys())}
g = f[sys.argv[2]]
def attr(obj, k):
    v = obj.attrs[k]
    t = obj.attrs.get_id(k).get_type()
    info = {'class': t.get_class(), 'size': t.get_size()}
    if t.get_class() == h5py.h5t.STRING:
        info['strpad'] = t.get_strpad()
        info['value'] = v.decode() if isinstance(v, bytes) else str(v)
    else:
        info['value'] = numpy.asarray(v).tolist()
    return info
out['group_attrs'] = {k: attr(g, k) for k in g.attrs}
out['datasets'] = {}
for name in g.keys():
    d = g[name]
    d_m = d.attrs['d_m']
    size = d.shape
    read_start = [abs(x) for x in d_m]
    read_end = [s - abs(x) for x, s in zip(d_m, size)]
    sl = tuple(slice(a, b) for a, b in zip(read_start, read_end))
    out['datasets'][name] = {'shape': list(d.shape), 'dtype': str(d.dtype), 'attrs': {k: attr(d, k) for k in d.attrs},
                             'all': d[()].ravel().tolist(), 'interior': d[()][sl].ravel().tolist()}
print(json.dumps(out))
"""

# writes a file as the reference's helper does (helperfunctions.py:80-127)
WRITE_LIKE_HELPER = r"""
import sys, json, h5py, numpy
fname, blockname, blocknumber = sys.argv[1], sys.argv[2], int(sys.argv[3])
spec = json.loads(sys.argv[4])
with h5py.File(fname, 'w') as hf:
    g1 = hf.create_group(blockname)
    for item in spec:
        ar = numpy.array(item['data'], dtype=numpy.float64).reshape(item['shape'])
        halos, npoints = item['halos'], item['npoints']
        ndim = len(npoints)
        g1.attrs.create('dims', [ndim], dtype='int32')
        g1.attrs.create('ops_type', u'ops_block', dtype='S9')
        g1.attrs.create('index', [blocknumber], dtype='int32')
        dset = g1.create_dataset(item['name'], data=ar)
        d_m = [halos[i][0] for i in range(ndim)]
        d_p = [halos[i][1] for i in range(ndim)]
        dset.attrs.create('d_p', d_p, dtype='int32')
        dset.attrs.create('d_m', d_m, dtype='int32')
        dset.attrs.create('dim', [1], dtype='int32')
        dset.attrs.create('ops_type', u'ops_dat', dtype='S10')
        dset.attrs.create('block_index', [blocknumber], dtype='int32')
        dset.attrs.create('base', [0 for i in range(ndim)], dtype='int32')
        dset.attrs.create('type', u'double', dtype='S15')
        dset.attrs.create('block', u'%s' % blockname, dtype='S25')
        dset.attrs.create('size', npoints, dtype='int32')
print(json.dumps({'ok': True}))
"""


def _block(ndim=2, number=0):
    from opensbli_amd.api import SimulationBlock
    return SimulationBlock(ndim, block_number=number)


def test_written_file_reads_like_the_reference_plot_scripts(tmp_path):
    hdf5 = _h5lib()
    f = str(tmp_path / "out.h5")
    rs = np.random.RandomState(1)
    # 3-D, asymmetric halos (3 below, 4 above in x; 2/2 in y; 0/1 in z): nx=5, ny=4, nz=3
    a = rs.standard_normal((3 + 0 + 1, 4 + 2 + 2, 5 + 3 + 4))
    b = rs.standard_normal((7 + 5 + 5, 6 + 5 + 5))
    hdf5.fetch_block_hdf5_file(f, "opensbliblock00", 0, 3, truncate=True)
    hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "rho_B0", a, [-3, -2, 0], [4, 2, 1])
    hdf5.fetch_block_hdf5_file(f, "opensbliblock01", 1, 2)
    hdf5.fetch_dat_hdf5_file(f, "opensbliblock01", 1, "x0_B1", b, [-5, -5], [5, 5])
    out = _conda(READ_LIKE_PLOT, f, "opensbliblock00")
    assert out["groups"] == ["opensbliblock00", "opensbliblock01"]
    ga = out["group_attrs"]
    assert ga["dims"]["value"] == [3] and ga["index"]["value"] == [0]
    assert ga["ops_type"] == {"class": 3, "size": 9, "strpad": 1, "value": "ops_block"}
    d = out["datasets"]["rho_B0"]
    assert d["shape"] == list(a.shape) and d["dtype"] == "float64"
    assert np.array_equal(np.array(d["all"]), a.ravel())
    at = {k: v["value"] for k, v in d["attrs"].items()}
    assert at == {"d_m": [-3, -2, 0], "d_p": [4, 2, 1], "dim": [1], "ops_type": "ops_dat", "block_index": [0],
                  "base": [0, 0, 0], "type": "double", "block": "opensbliblock00", "size": [5, 4, 3]}
    sizes = {k: (v["class"], v["size"]) for k, v in d["attrs"].items()}
    assert sizes["ops_type"] == (3, 10) and sizes["type"] == (3, 15) and sizes["block"] == (3, 25)
    assert all(sizes[k] == (0, 4) for k in ("d_m", "d_p", "dim", "block_index", "base", "size"))
    # the plot scripts pair d_m with the dataset's axes in order; with equal depths per axis their
    # interior is the library's own interior
    out1 = _conda(READ_LIKE_PLOT, f, "opensbliblock01")
    assert np.array_equal(np.array(out1["datasets"]["x0_B1"]["interior"]), b[5:-5, 5:-5].ravel())
    assert np.array_equal(hdf5.read_interior(f, "opensbliblock01", "x0_B1"), b[5:-5, 5:-5])
    assert np.array_equal(hdf5.read_interior(f, "opensbliblock00", "rho_B0"), a[0:3, 2:6, 3:8])


def test_reads_files_written_like_the_reference_helper(tmp_path):
    hdf5 = _h5lib()
    f = str(tmp_path / "data.h5")
    rs = np.random.RandomState(2)
    noise = rs.rand(6 + 10, 6 + 10)
    coords = rs.rand(9 + 4, 7 + 6)
    spec = [{"name": "random_nums_B0", "data": noise.ravel().tolist(), "shape": list(noise.shape),
             "halos": [[-5, 5], [-5, 5]], "npoints": [6, 6]},
            {"name": "x1_B0", "data": coords.ravel().tolist(), "shape": list(coords.shape),
             "halos": [[-3, 3], [-2, 2]], "npoints": [7, 9]}]
    _conda(WRITE_LIKE_HELPER, f, "opensbliblock00", "0", json.dumps(spec))
    a, info = hdf5.decl_dat_hdf5(f, "opensbliblock00", "random_nums_B0")
    assert info == {"ndim": 2, "size": [6, 6], "d_m": [-5, -5], "d_p": [5, 5]}
    assert np.array_equal(a, noise)
    assert np.array_equal(hdf5.read_interior(f, "opensbliblock00", "random_nums_B0"), noise[5:-5, 5:-5])
    info = hdf5.dat_info(f, "opensbliblock00", "x1_B0")
    assert info == {"ndim": 2, "size": [7, 9], "d_m": [-3, -2], "d_p": [3, 2]}
    assert np.array_equal(hdf5.read_interior(f, "opensbliblock00", "x1_B0"), coords[2:-2, 3:-3])
    assert sorted(hdf5.list_names(f, "opensbliblock00")) == ["random_nums_B0", "x1_B0"]


def test_output_hdf5_metadata_matches_the_reference_helper(tmp_path):
    """h5dump of output_hdf5's file equals h5dump of the file the reference's
    helper writes with h5py for the same array (types, sizes, padding, values)."""
    _h5lib()
    if not os.path.exists(H5DUMP):
        pytest.skip("no h5dump in this image")
    from opensbli_amd.api import output_hdf5
    rs = np.random.RandomState(3)
    noise = rs.rand(8 + 10, 8 + 10)
    ours, theirs = str(tmp_path / "ours.h5"), str(tmp_path / "theirs.h5")
    output_hdf5(noise, "random_nums", [(-5, 5), (-5, 5)], [8, 8], _block(), filename=ours)
    spec = [{"name": "random_nums_B0", "data": noise.ravel().tolist(), "shape": list(noise.shape),
             "halos": [[-5, 5], [-5, 5]], "npoints": [8, 8]}]
    _conda(WRITE_LIKE_HELPER, theirs, "opensbliblock00", "0", json.dumps(spec))

    def dump(p):
        txt = subprocess.run([H5DUMP, p], capture_output=True, text=True, check=True).stdout
        return txt.split("\n", 1)[1]
    assert dump(ours) == dump(theirs)


def test_errors(tmp_path):
    hdf5 = _h5lib()
    f = str(tmp_path / "e.h5")
    with pytest.raises(hdf5.H5MissingError, match="no such file"):
        hdf5.dat_info(f, "opensbliblock00", "rho_B0")
    with pytest.raises(hdf5.H5MissingError, match="fetch_block"):
        hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "rho_B0", np.zeros((4, 4)), [-1, -1], [1, 1])
    hdf5.fetch_block_hdf5_file(f, "opensbliblock00", 0, 2, truncate=True)
    with pytest.raises(hdf5.H5MissingError, match="no block group"):
        hdf5.dat_info(f, "opensbliblock01", "rho_B0")
    with pytest.raises(hdf5.H5MissingError, match="no dataset"):
        hdf5.dat_info(f, "opensbliblock00", "rho_B0")
    with pytest.raises(hdf5.H5Error, match="must be <= 0"):
        hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "rho_B0", np.zeros((4, 4)), [1, -1], [1, 1])
    with pytest.raises(hdf5.H5Error, match="contains '/'"):
        hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "a/b", np.zeros((4, 4)), [-1, -1], [1, 1])
    with pytest.raises(hdf5.H5Error, match="longer than 25"):
        hdf5.fetch_block_hdf5_file(f, "b" * 26, 0, 2)
        hdf5.fetch_dat_hdf5_file(f, "b" * 26, 0, "rho_B0", np.zeros((4, 4)), [-1, -1], [1, 1])
    with pytest.raises(hdf5.H5Error, match="not an HDF5 file"):
        p = tmp_path / "text.h5"
        p.write_text("not hdf5")
        hdf5.dat_info(str(p), "opensbliblock00", "rho_B0")
    # a dataset whose attributes disagree with its extent
    hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "rho_B0", np.zeros((6, 6)), [-1, -1], [1, 1])
    spec = [{"name": "bad_B0", "data": [0.0] * 36, "shape": [6, 6], "halos": [[-1, 1], [-1, 1]], "npoints": [5, 4]}]
    g = str(tmp_path / "bad.h5")
    _conda(WRITE_LIKE_HELPER, g, "opensbliblock00", "0", json.dumps(spec))
    with pytest.raises(hdf5.H5Error, match="does not match its extent"):
        hdf5.dat_info(g, "opensbliblock00", "bad_B0")
    # replacing a dataset with another shape, and appending to an existing file
    hdf5.fetch_dat_hdf5_file(f, "opensbliblock00", 0, "rho_B0", np.ones((7, 5)), [-1, -2], [1, 2])
    hdf5.fetch_block_hdf5_file(f, "opensbliblock00", 0, 2)
    assert hdf5.dat_info(f, "opensbliblock00", "rho_B0")["size"] == [3, 3]
    assert np.array_equal(hdf5.read_interior(f, "opensbliblock00", "rho_B0"), np.ones((3, 3)))


def test_iohdf5_surface():
    from opensbli_amd.api import DataObject, iohdf5
    w = iohdf5()
    assert w.iotype == "write" and w.file_name() == "opensbli_output.h5"
    s = iohdf5(save_every=100, **{"iotype": "Write"})
    assert s.file_name(300) == "opensbli_output_000300.h5" and s.file_name() == "opensbli_output.h5"
    assert [type(p).__name__ for p in s.algorithm_place] == ["InTheSimulation", "AfterSimulationEnds"]
    r = iohdf5(save_every=10, **{"iotype": "Read", "filename": "Restart.H5"})
    assert [type(p).__name__ for p in r.algorithm_place] == ["BeforeSimulationStarts"]
    assert r.read_file_name() == "restart.h5"  # keyword values are lower-cased (io_hdf5.py:26-29)
    assert iohdf5(**{"iotype": "read"}).read_file_name() == "data.h5"
    with pytest.raises(ValueError):
        iohdf5(**{"iotype": "append"})
    assert iohdf5(**{"iotype": "write", "name": "final"}).file_name() == "final.h5"
    with pytest.raises(ValueError):
        iohdf5(**{"iotype": "write", "name": "final.dat"}).file_name()
    with pytest.raises(ValueError, match="dynamic fname"):
        iohdf5(save_every=5, **{"iotype": "write", "name": "final"}).file_name(5)
    w.add_arrays([DataObject("rho"), [DataObject("rhoE"), DataObject("T")]])
    with pytest.raises(ValueError, match="'T' added to the HDF5 class are not defined"):
        w.check_datasets({"rho", "rhoE"})


def test_output_hdf5_argument_checks(tmp_path):
    _h5lib()
    from opensbli_amd.api import output_hdf5
    blk = _block()
    a = np.zeros((14, 14))
    with pytest.raises(ValueError, match="Two halos should be provided"):
        output_hdf5(a, "n", [(-2, 2)] * 3, [10, 10], blk, filename=str(tmp_path / "a.h5"))
    with pytest.raises(ValueError, match="should be of size 2"):
        output_hdf5(a, "n", [(-2, 2, 0), (-2, 2, 0)], [10, 10], blk, filename=str(tmp_path / "a.h5"))
    with pytest.raises(ValueError, match="is not npoints"):
        output_hdf5(a, "n", [(-2, 2), (-2, 2)], [10, 9], blk, filename=str(tmp_path / "a.h5"))
    # 3-D: one (low, high) pair per direction
    b = np.zeros((5 + 2, 4 + 2, 3 + 2))
    output_hdf5(b, "n", [(-1, 1)] * 3, [3, 4, 5], _block(3), filename=str(tmp_path / "b.h5"))
    from opensbli_amd import hdf5
    assert hdf5.dat_info(str(tmp_path / "b.h5"), "opensbliblock00", "n_B0")["size"] == [3, 4, 5]


def _kh_ic(noise, n):
    """The KH app's initial condition (kelvin_helmholtz.py:59-73), restated
    with numpy on the interior points, from the interior of the noise array."""
    j = np.arange(n)[:, None] * np.ones((1, n))
    x1 = -0.5 + j / n
    inner = (x1 >= -0.25) & (x1 <= 0.25)
    d = np.where(inner, 2.0, 1.0)
    u0 = np.where(inner, 0.5 + 0.01 * (noise - 0.5), -0.5)
    u1 = np.where(inner, 0.01 * (noise - 0.5), 0.0)
    p = 2.5
    return np.stack([d, d * u0, d * u1, p / 0.4 + 0.5 * d * (u0 ** 2 + u1 ** 2)])


def test_kh_app_reads_its_noise_from_hdf5(tmp_path, monkeypatch):
    _h5lib()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd import hdf5
    from opensbli_amd.apps.kelvin_helmholtz import build
    n = 24
    prog = build(n, niter=10, save_every=5, seed=4)
    assert os.path.exists(tmp_path / "data.h5")
    noise = hdf5.read_interior("data.h5", "opensbliblock00", "random_nums_B0")
    assert np.array_equal(noise, np.random.RandomState(4).rand(n + 10, n + 10)[5:-5, 5:-5])
    p = prog.problem()
    q0 = prog.initial_state(p)
    ref = _kh_ic(noise, n)
    assert np.abs(q0 - ref).max() <= 1e-15 * np.abs(ref).max()
    assert prog._defined_arrays(2) >= {"rho", "rhou0", "rhou1", "rhoE", "x0", "x1", "random_nums"}
    # coordinates over the grid widened by halos, as the initialisation kernel covers them
    env, shape = prog._env(p, [-2, -3], [2, 1])
    assert shape == (n + 4, n + 4)
    out = {}
    prog._evaluate(prog._init_equations(), env, shape, out)
    assert np.allclose(out["x0"][0], -0.5 + np.arange(-2, n + 2) / n, rtol=0, atol=1e-15)
    assert np.allclose(out["x1"][:, 0], -0.5 + np.arange(-3, n + 1) / n, rtol=0, atol=1e-15)
    with pytest.raises(ValueError, match="halos -5/5"):
        prog._env(p, [-6, 0], [0, 0])


def test_restart_state_from_hdf5(tmp_path, monkeypatch):
    """An iohdf5 read of the conservative arrays with no initial condition
    gives them as the initial state."""
    _h5lib()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd import hdf5
    from opensbli_amd.api import (HIPC, ConstituentRelations, DataObject, EinsteinEquation, LLFWeno, PeriodicBC,
                                  RoeAverage, RungeKutta, SimulationBlock, SimulationEquations,
                                  TraditionalAlgorithmRK, iohdf5, substitute_simulation_parameters)
    n = 12
    sc = "**{'scheme':'Weno'}"
    ee = EinsteinEquation()
    sim, cr = SimulationEquations(), ConstituentRelations()
    for t in ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc,
              "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc,
              "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s))" % sc]:
        sim.add_equations(ee.expand(t, 2, "x", [], ["gama", "Minf"]))
    for t in ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
              "Eq(a, (gama*p/rho)**0.5)"]:
        cr.add_equations(ee.expand(t, 2, "x", [], ["gama", "Minf"]))
    block = SimulationBlock(2, block_number=0)
    llf, rk = LLFWeno(5, formulation="Z", averaging=RoeAverage([0, 1])), RungeKutta(3)
    block.set_block_boundaries([PeriodicBC(d, s) for d in range(2) for s in (0, 1)])
    block.set_discretisation_schemes({llf.name: llf, rk.name: rk})
    rd = iohdf5(**{"iotype": "Read", "filename": "restart.h5"})
    rd.add_arrays(sim.time_advance_arrays)
    block.setio([rd])
    block.set_equations([cr, sim])
    block.discretise()
    prog = HIPC(TraditionalAlgorithmRK(block))
    substitute_simulation_parameters(["gama", "Minf", "dt", "niter", "block0np0", "block0np1", "Delta0block0",
                                      "Delta1block0"],
                                     ["1.4", "2.0", "1e-4", "3", str(n), str(n), "1.0/block0np0", "1.0/block0np1"])
    rs = np.random.RandomState(5)
    q = 1.0 + 0.1 * rs.rand(4, n, n)
    hdf5.fetch_block_hdf5_file("restart.h5", "opensbliblock00", 0, 2, truncate=True)
    for m, nm in enumerate(["rho", "rhou0", "rhou1", "rhoE"]):
        padded = np.pad(q[m], 4, mode="wrap")
        hdf5.fetch_dat_hdf5_file("restart.h5", "opensbliblock00", 0, nm + "_B0", padded, [-4, -4], [4, 4])
    q0 = prog.initial_state(prog.problem())
    assert np.array_equal(q0, q)
    os.remove("restart.h5")
    with pytest.raises(KeyError, match="no such file"):
        prog.initial_state(prog.problem())
