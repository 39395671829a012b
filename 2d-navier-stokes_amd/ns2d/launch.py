"""Run the solver from the reference's .ini job files.

Reads the [SYSTEM], [SOLVER], [TIME], [DIRECTORIES] and [JOB] sections with
the keys and defaults of run_solver.py (run_solver.py:75-214) and builds one
solver command per combination of the listed grid sizes, end times, time
steps, initial conditions, viscosities, hyperviscosity flags, CFL numbers and
tags, with the argument layout of run_solver.py:232-245.  The commands call
the single-GPU ns2d_solver directly instead of ``mpirun -n <procs>`` and the
reference executable; ``-R <procs>`` asks it for the initial condition that
many MPI ranks would have drawn (the reference's random initial conditions
depend on the rank count).  Bracketed lists such as ``[1.0, 2.0]`` are accepted
for every list-valued key.

    python -m ns2d.launch -i InitFiles/TaylorGreen/tg.ini [--cmdonly]
"""
from __future__ import annotations

import getopt
import itertools
import os
import subprocess
import sys
from configparser import ConfigParser
from dataclasses import dataclass, field
from typing import List

from .host import SOLVER_EXE


def _list(text: str) -> List[str]:
    return [t.strip().strip('"') for t in text.strip().lstrip("[").rstrip("]").split(",") if t.strip()]


def _true(text: str) -> bool:
    return text.strip() == "True"


@dataclass
class Job:
    """The run_solver.py parameters (defaults at run_solver.py:75-118)."""
    Nx: List[int] = field(default_factory=list)
    Ny: List[int] = field(default_factory=list)
    nu: List[float] = field(default_factory=list)
    hyper_visc: List[bool] = field(default_factory=list)
    ic: List[str] = field(default_factory=list)
    T: List[float] = field(default_factory=list)
    dt: List[float] = field(default_factory=list)
    cfl: List[float] = field(default_factory=list)
    solver_tag: List[str] = field(default_factory=list)
    ekmn_alpha: float = 1.0
    hypervisc_pow: float = 2.0
    ekmn_hypo_diff: int = 0
    ekmn_hypo_pow: float = -2.0
    t0: float = 0.0
    step_type: int = 1
    cfl_cond: int = 1
    trans_iters: int = 1
    forcing: str = "NONE"
    force_k: int = 0
    force_scale: float = 1.0
    save_every: int = 2
    output_dir: str = "./Data/Tmp/"
    input_dir: str = "NONE"
    file_only_mode: bool = False
    call_solver: bool = True
    num_solver_job_threads: int = 1
    solver_procs: int = 4


def read_ini(path: str) -> Job:
    """The parameter parse of run_solver.py:119-240."""
    if not os.path.isfile(path):
        raise FileNotFoundError(f"[ERROR] ---> File Does not exist, double check input file path: {path}")
    cp = ConfigParser()
    cp.read(path)
    j = Job()
    if cp.has_section("SYSTEM"):
        s = cp["SYSTEM"]
        j.Nx = [int(v) for v in _list(s.get("nx", ""))]
        j.Ny = [int(v) for v in _list(s.get("ny", ""))]
        j.nu = [float(v) for v in _list(s.get("viscosity", ""))]
        j.hyper_visc = [v == "True" for v in _list(s.get("hyperviscosity", ""))]
        if "drag_coefficient" in s:
            j.ekmn_alpha = float(s["drag_coefficient"])
        if "hypo_diffusion" in s:
            j.ekmn_hypo_diff = int(_true(s["hypo_diffusion"]))
        if "hyperviscosity_pow" in s:
            j.hypervisc_pow = float(s["hyperviscosity_pow"])
        if "hypo_diffusion_pow" in s:
            j.ekmn_hypo_pow = float(s["hypo_diffusion_pow"])
    if cp.has_section("SOLVER"):
        s = cp["SOLVER"]
        j.ic = _list(s.get("initial_condition", ""))
        if "forcing" in s:
            j.forcing = s["forcing"].strip()
        if "forcing_wavenumber" in s:
            j.force_k = int(float(s["forcing_wavenumber"]))
        if "forcing_scale" in s:
            j.force_scale = float(s["forcing_scale"])
        if "save_data_every" in s:
            j.save_every = int(s["save_data_every"])
    if cp.has_section("TIME"):
        s = cp["TIME"]
        j.T = [float(v) for v in _list(s.get("end_time", ""))]
        j.dt = [float(v) for v in _list(s.get("timestep", ""))]
        j.cfl = [float(v) for v in _list(s.get("cfl", ""))]
        if "start_time" in s:
            j.t0 = float(s["start_time"])
        if "cfl_cond" in s:
            j.cfl_cond = int(_true(s["cfl_cond"]))
        if "trans_iters" in s:
            j.trans_iters = int(_true(s["trans_iters"]))
        if "adaptive_step_type" in s:
            j.step_type = int(_true(s["adaptive_step_type"]))
    if cp.has_section("DIRECTORIES"):
        s = cp["DIRECTORIES"]
        if "solver_input_dir" in s:
            j.input_dir = s["solver_input_dir"].strip()
        if "solver_output_dir" in s:
            j.output_dir = s["solver_output_dir"].strip()
        j.solver_tag = _list(s.get("solver_tag", ""))
        if "solver_file_only_mode" in s:
            j.file_only_mode = s["solver_file_only_mode"].strip().lower() in ("true", "1", "yes", "on")
    if cp.has_section("JOB"):
        s = cp["JOB"]
        if "call_solver" in s:
            j.call_solver = s["call_solver"].strip().lower() in ("true", "1", "yes", "on")
        if "num_solver_job_threads" in s:
            j.num_solver_job_threads = int(s["num_solver_job_threads"])
        if "solver_procs" in s:
            j.solver_procs = int(s["solver_procs"])
    return j


def solver_commands(j: Job, executable: str = SOLVER_EXE) -> List[List[str]]:
    """One argument list per parameter combination, in run_solver.py's order
    (sizes, end time, time step, initial condition, viscosity, hyperviscosity
    flag, CFL, tag), each ending in ``-R solver_procs`` (run_solver.py:270
    runs ``mpirun -n solver_procs``)."""
    cmds = []
    for (nx, ny), t, h, u0, v, hype, c, tag in itertools.product(
            list(zip(j.Nx, j.Ny)), j.T, j.dt, j.ic, j.nu, j.hyper_visc, j.cfl, j.solver_tag):
        args = ("-o {} -n {} -n {} -s {:3.5f} -e {:3.5f} -T {} -c {} -c {:1.6f} -h {:1.6f} -h {} "
                "-v {:1.10f} -v {} -v {:1.1f} -d {:1.6f} -d {} -d {:1.1f} -i {} -t {} -f {} -f {} -f {} -p {}").format(
            j.output_dir, nx, ny, j.t0, t, j.trans_iters, j.cfl_cond, c, h, j.step_type, v, int(hype),
            j.hypervisc_pow, j.ekmn_alpha, j.ekmn_hypo_diff, j.ekmn_hypo_pow, u0, tag, j.forcing, j.force_k,
            j.force_scale, j.save_every)
        cmds.append([executable] + args.split() + ["-R", str(j.solver_procs)])
    return cmds


def main(argv: List[str]) -> int:
    try:
        opts, _ = getopt.getopt(argv, "i:c:", ["cmdonly"])
    except getopt.GetoptError as e:
        print("[ERROR] ---> Incorrect Command Line Arguements.\n" + str(e))
        return 1
    ini, cmd_only = None, False
    for o, a in opts:
        if o == "-i":
            ini = a
        elif o == "--cmdonly":
            cmd_only = True
    if ini is None:
        print("usage: python -m ns2d.launch -i <file.ini> [--cmdonly]")
        return 1
    j = read_ini(ini)
    cmds = solver_commands(j)
    if cmd_only or not j.call_solver:
        for c in cmds:
            print(" ".join(c) + "\n")
        return 0
    rc = 0
    # the jobs share the one GPU: run them one after another
    for c in cmds:
        print("Executing the following command:\n\t" + " ".join(c), flush=True)
        rc = subprocess.call(c) or rc
    return rc


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
