"""The documented binding of the drop-in, built and run: the
reference's own program -- its main.c, readpar.c, gridsetup.c, domain.c,
report.c, misc.c (set-up), the Hydro/HLLD/Boundary/Initial/Planets modules --
compiled from /root/reference without timestep.c, onestep.c and analysis.c,
with main.c's getmindt wrapped (-Wl,--wrap=getmindt), linked to this
project's host layer, which takes the reference-built domain onto the device
at the first getmindt and keeps the host cells current whenever the
reference's possiblyOutput will read them.  The same program built whole
(the reference alone) runs the same in.par; the final checkpoint (written by
an Output plugin of this test through the reference's possiblyOutput) and
report.dat agree bit for bit, the device side being the engine's host build
(tests/cpu_engine), which computes what the GPU library computes.

Also: struct domain, cell, face, planet and param_list of include/disco/
paul.h have the reference paul.h's layout, offset by offset."""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from disco_amd import configs, state_io

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF = "/root/reference"
PLUG = os.path.join(REPO, "disco_amd", "plugins")
sys.path.insert(0, os.path.join(HERE, "cpu_engine"))
import build as cpu_build  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not mounted")

CORE = ["readpar.c", "timestep.c", "onestep.c", "riemann.c", "mpisetup.c", "gridsetup.c",
        "domain.c", "misc.c", "faces_alt.c", "exchange.c", "plm.c", "report.c", "profiler.c",
        "planet.c", "omega.c", "analysis.c", "bfields.c", "rotframe.c",
        "boundary_functions.c", "geometry_functions.c", "noise.c", "sink.c",
        "Restart/no_restart.c", "Geometry/cylindrical.c", "Hydro/Frame/normal.c", "main.c"]
# the objects the binding leaves out: the step's own files, whose entry
# points the host layer supplies
REPLACED = {"timestep.c", "onestep.c", "analysis.c"}

CASES = {
    "euler_yee": dict(hydro="euler", hlld="hlld", metric="flat_rpz", boundary=f"{REF}/Boundary/polar.c",
                      planets=f"{REF}/Planets/none.c", analysis="diag_euler",
                      initial=f"{PLUG}/Initial/yee_vortex.c", num_c=5, ct=0, extra=[]),
    "mhd_kep": dict(hydro="mhd", hlld="hlld", metric="flat_rpz",
                    boundary=f"{PLUG}/Boundary/fixed_r_outflow_z.c",
                    planets=f"{PLUG}/Planets/central.c", analysis="diag_mhdmri",
                    initial=f"{PLUG}/Initial/kep_mhd_disk.c", num_c=8, ct=2,
                    extra=[os.path.join(REPO, "oracle", "ref", "visc_source_noop.c")]),
    "gr_disk": dict(hydro="grmhd2", hlld="hlld_grmhd", metric="schw_ks_eq_rpz",
                    boundary=f"{PLUG}/Boundary/fixed_r_outflow_z.c",
                    planets=f"{REF}/Planets/none.c", analysis="diag_euler",
                    initial=f"{PLUG}/Initial/gr_disk.c", num_c=8, ct=2, extra=[]),
}


def _defs(c, output):
    name = lambda p: os.path.splitext(os.path.basename(p))[0]  # noqa: E731
    return ["-DUSE_MPI=0", '-DGIT_VERSION="binding-test"', f'-DINITIAL="{name(c["initial"])}"',
            f'-DHYDRO="{c["hydro"]}"', '-DGEOMETRY="cylindrical"',
            f'-DBOUNDARY="{name(c["boundary"])}"', f'-DOUTPUT="{output}"', '-DRESTART="no_restart"',
            f'-DPLANETS="{name(c["planets"])}"', f'-DHLLD="{c["hlld"]}"',
            f'-DANALYSIS="{c["analysis"]}"', f'-DMETRIC="{c["metric"]}"', '-DFRAME="normal"',
            f"-DNUM_C={c['num_c']}", "-DNUM_N=0", f"-DCT_MODE={c['ct']}"]


H5 = "/opt/conda"   # HDF5 (headers and library) of the reference's checkpoints


def _build(cname, d, hdf5=False):
    """(whole reference program, bound program) of one module selection;
    with hdf5 the reference's own Output/h5out.c and Restart/h5in.c modules,
    otherwise this test's DSTATE01 Output plugin."""
    c = CASES[cname]
    output, restart = ("h5out", "h5in") if hdf5 else ("dstate_output", "no_restart")
    defs = [f for f in _defs(c, output) if not f.startswith("-DRESTART")]
    cflags = ["gcc", "-O3", "-ffp-contract=off", "-w", f"-I{REF}",
              f"-I{os.path.join(REPO, 'disco_amd', 'csrc')}", *defs, f'-DRESTART="{restart}"']
    libs = ["-lm"]
    if hdf5:
        cflags.append(f"-I{H5}/include")
        # by path, and the system's libraries first for what the GPU runtime
        # needs (HDF5's directory holds an older libstdc++)
        libs = [f"{H5}/lib/libhdf5.so", "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu",
                f"-Wl,-rpath,{H5}/lib", "-lm"]
    out_src = (f"{REF}/Output/h5out.c" if hdf5
               else os.path.join(HERE, "binding", "dstate_output.c"))
    plugins = [c["initial"], c["boundary"], c["planets"], f"{REF}/Diagnostics/{c['analysis']}.c",
               f"{REF}/Hydro/Metric/{c['metric']}.c", f"{REF}/Hydro/{c['hydro']}.c",
               f"{REF}/{c['hlld']}.c", out_src, *c["extra"]]
    core = [f for f in CORE if f != "Restart/no_restart.c"] + [f"Restart/{restart}.c"
                                                                if hdf5 else "Restart/no_restart.c"]
    whole = os.path.join(d, "disco_reference")
    subprocess.run(cflags + [f"{REF}/{f}" for f in core] + plugins + libs + ["-o", whole],
                   check=True)
    # the host layer as the binding links it: disco_host.c alone (the
    # plugins come from the program), on the engine's host build
    engine = cpu_build.build()
    from disco_amd.build import HOST_CONFIGS, host_defines
    hc = HOST_CONFIGS[cname]
    lib = os.path.join(d, f"libdisco_host_{cname}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-parameter",
                    *host_defines(hc), f"-I{os.path.join(REPO, 'include', 'disco')}",
                    f"-I{os.path.join(REPO, 'include')}",
                    os.path.join(REPO, "disco_amd", "host", "disco_host.c"), "-o", lib,
                    f"-L{os.path.dirname(engine)}", "-ldh_cpu", "-lm",
                    f"-Wl,-rpath,{os.path.dirname(engine)}"], check=True)
    bound = os.path.join(d, "disco_bound")
    kept = [f"{REF}/{f}" for f in core if f not in REPLACED]
    subprocess.run(cflags + kept + plugins + ["-rdynamic", "-Wl,--wrap=getmindt", f"-L{d}",
                                              f"-ldisco_host_{cname}", f"-Wl,-rpath,{d}", *libs,
                                              "-o", bound], check=True)
    return whole, bound


@pytest.mark.parametrize("cname,over,steps", [
    ("euler_yee", {"Num_R": 24}, 8),
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6}, 6),
    ("gr_disk", {"Num_R": 16, "Num_Z": 4, "Riemann_Solver": 2, "CFL": 0.2}, 5),
])
def test_reference_main_runs_on_the_drop_in(cname, over, steps):
    cfg = configs.get(cname)
    d = tempfile.mkdtemp(prefix="bind_")
    whole, bound = _build(cname, d)
    # a fixed step below the CFL step, T_End just short of `steps` of them
    # (check_dt trims the last), five reports on the way
    h = 1e-3 if cname != "gr_disk" else 2e-2
    par = cfg.par_dict(**over, Max_DT=h, T_End=steps * h * (1 - 1e-9), Num_Reports=5,
                       Num_Checkpoints=0)
    runs = {}
    for tag, exe in (("whole", whole), ("bound", bound)):
        wd = os.path.join(d, tag)
        state_io.write_par(wd, par)
        r = subprocess.run([exe], cwd=wd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (tag, r.stdout[-2000:], r.stderr[-2000:])
        runs[tag] = wd
    a = state_io.read_state(os.path.join(runs["whole"], "output.dstate"))
    b = state_io.read_state(os.path.join(runs["bound"], "output.dstate"))
    assert a.step == b.step == steps and a.t == b.t
    for key in ("prim", "piph", "dphi", "wiph", "phi_flux", "planets"):
        x, y = getattr(a, key), getattr(b, key)
        assert np.array_equal(x, y), (key, np.abs(x - y).max())
    # conserved variables of real cells (ghost cells' are not state: no step
    # reads them, and the device keeps only their primitives)
    m = a.real_mask()
    assert np.array_equal(a.cons[:, m], b.cons[:, m])
    rep_a = open(os.path.join(runs["whole"], "report.dat")).read()
    rep_b = open(os.path.join(runs["bound"], "report.dat")).read()
    assert rep_a == rep_b and rep_a.count("\n") >= 3


def _layout(header_dir, extra_flags):
    fields = {
        "struct domain": ["theCells", "Np", "Nr", "NgZb", "N_ftracks_z", "Npl", "r_jph", "phi_max",
                          "fIndex_z", "Wallt_init", "dim_size", "right_rank", "theParList",
                          "num_tools", "theTools", "t", "count_steps", "t_fin", "nrpt", "N_chk",
                          "final_step", "check_plz"],
        "struct cell": ["prim", "cons", "RKcons", "gradr", "gradp", "gradz", "piph", "dphi",
                        "wiph", "E", "B", "E_phi", "Phi", "RK_Phi", "tempDoub", "real"],
        "struct planet": ["r", "phi", "M", "omega", "vr", "RK_r", "RK_phi", "RK_M", "RK_omega",
                          "RK_vr", "eps", "Fr", "Fp", "type"],
        "struct param_list": ["t_min", "Num_R", "aspect", "rmin", "phimax", "NoBC_Zmax", "LogRadius",
                              "Mesh_Motion", "CFL", "maxDT", "Density_Floor", "Exact_Mesh_Omega",
                              "RotFrame", "Adiabatic_Index", "isothermal_flag", "Disk_Mach",
                              "grav2D", "alpha_flag", "restart_flag", "CT", "metricPar0",
                              "metricPar4", "initPar0", "initPar8", "noiseType", "noiseRel",
                              "sinkType", "nozzlePar4"],
    }
    src = ['#include <stddef.h>', '#include <stdio.h>', '#include "paul.h"', "int main(void) {"]
    for st, fs in fields.items():
        src.append(f'printf("%zu\\n", sizeof({st}));')
        src += [f'printf("%zu\\n", offsetof({st}, {f}));' for f in fs]
    src.append("return 0; }")
    d = tempfile.mkdtemp(prefix="paul_")
    c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
    with open(c, "w") as f:
        f.write("\n".join(src))
    subprocess.run(["gcc", f"-I{header_dir}", *extra_flags, c, "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()


@pytest.mark.parametrize("sizes", [("5", "0", "0"), ("5", "1", "0"), ("8", "0", "2"),
                                   ("8", "1", "2"), ("8", "0", "1")])
def test_plugin_abi_structs_have_the_reference_layout(sizes):
    flags = [f"-DNUM_C={sizes[0]}", f"-DNUM_N={sizes[1]}", f"-DCT_MODE={sizes[2]}"]
    ours = _layout(os.path.join(REPO, "include", "disco"), flags)
    ref = _layout(REF, ["-DUSE_MPI=0", *flags])
    assert ours == ref


def _cells(h5, d):
    """(T, Grid/Np[Nz][Nr], Data/Cells) of a checkpoint in the reference's
    layout (cells annulus by annulus, k outer, j inner)."""
    reader = os.path.join(d, "h5cells")
    if not os.path.exists(reader):
        subprocess.run(["gcc", "-O2", f"-I{H5}/include", os.path.join(HERE, "binding", "h5cells.c"),
                        f"-L{H5}/lib", "-lhdf5", f"-Wl,-rpath,{H5}/lib", "-o", reader], check=True)
    raw = h5 + ".bin"
    subprocess.run([reader, h5, raw], check=True)
    buf = open(raw, "rb").read()
    T = np.frombuffer(buf, np.float64, 1, 0)[0]
    nn = np.frombuffer(buf, np.int64, 2, 8)
    o = 24
    Np = np.frombuffer(buf, np.int64, int(nn[0] * nn[1]), o).reshape(nn)
    o += Np.nbytes
    n = np.frombuffer(buf, np.int64, 2, o)
    return T, Np, np.frombuffer(buf, np.float64, int(n[0] * n[1]), o + 16).reshape(n)


@pytest.mark.skipif(not os.path.exists(f"{H5}/include/hdf5.h"), reason="no HDF5 headers")
@pytest.mark.parametrize("cname,over,n1,n2", [
    ("euler_yee", {"Num_R": 24}, 4, 4),
    ("mhd_kep", {"Num_R": 12, "Num_Z": 6}, 3, 3),
])
def test_checkpoints_in_the_reference_layout_restart_across(cname, over, n1, n2):
    """Checkpoints in the reference's HDF5 layout (Output/h5out.c) and
    restarts from them (Restart/h5in.c), across the two programs: the drop-in
    continues a run from the reference's checkpoint, and the reference
    continues one from the drop-in's, each ending where the reference's own
    split run ends, bit for bit."""
    cfg = configs.get(cname)
    d = tempfile.mkdtemp(prefix="bindh5_")
    progs = dict(zip(("whole", "bound"), _build(cname, d, hdf5=True)))
    h = 1e-3
    first = cfg.par_dict(**over, Max_DT=h, T_End=n1 * h * (1 - 1e-9), Num_Reports=0,
                         Num_Checkpoints=0)
    second = dict(first, T_End=(n1 + n2) * h * (1 - 1e-9), Restart=1)
    ends = {}
    for a in ("whole", "bound"):           # who writes the checkpoint
        wa = os.path.join(d, f"{a}_first")
        state_io.write_par(wa, first)
        r = subprocess.run([progs[a]], cwd=wa, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (a, r.stderr[-2000:])
        for b in ("whole", "bound"):       # who continues from it
            wb = os.path.join(d, f"{a}_then_{b}")
            state_io.write_par(wb, second)
            os.link(os.path.join(wa, "output.h5"), os.path.join(wb, "input.h5"))
            r = subprocess.run([progs[b]], cwd=wb, capture_output=True, text=True, timeout=600)
            assert r.returncode == 0, (a, b, r.stderr[-2000:])
            ends[(a, b)] = _cells(os.path.join(wb, "output.h5"), d)
    T0, Np, ref = ends[("whole", "whole")]
    assert abs(T0 - (n1 + n2) * h) < 1e-9 and ref.shape[1] == cfg.nq + cfg.nf + 1
    # real cells: the ghost annuli (two on every side with a boundary) carry
    # boundary data; the reference's own restart leaves parts of its ghost
    # cells' edge and flux records unset, so they are compared on real cells
    nz, nr = Np.shape
    gr = [0 if first.get(k) else 2 for k in ("NoBC_Rmin", "NoBC_Rmax")]
    gz = [0 if first.get(k) or nz == 1 else 2 for k in ("NoBC_Zmin", "NoBC_Zmax")]
    real = np.zeros((nz, nr), bool)
    real[gz[0]:nz - gz[1], gr[0]:nr - gr[1]] = True
    m = np.repeat(real.ravel(), Np.ravel())
    for k, (T, Np_k, cells) in ends.items():
        assert T == T0 and np.array_equal(Np_k, Np), k
        assert np.array_equal(cells[m], ref[m]), (k, np.abs(cells[m] - ref[m]).max())
