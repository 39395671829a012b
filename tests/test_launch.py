"""CPU tests of the .ini job launcher (run_solver.py's parameter parse and
command layout) feeding the driver's command-line parser."""
import os

from ns2d import host, launch

INI = os.path.join(os.path.dirname(__file__), "golden", "job_example.ini")


def test_commands_follow_run_solver_layout():
    j = launch.read_ini(INI)
    cmds = launch.solver_commands(j, executable="ns2d_solver")
    # sizes zip pairwise, then the product over viscosities (run_solver.py:232-245)
    assert len(cmds) == 4
    assert " ".join(cmds[0]) == (
        "ns2d_solver -o ./Data/Test/ -n 64 -n 64 -s 0.00000 -e 1.00000 -T 0 -c 1 -c 0.900000 -h 0.001000 -h 0 "
        "-v 0.0010000000 -v 0 -v 2.0 -d 0.000000 -d 1 -d 0.0 -i TG_VORT -t TG-Test -f NONE -f 0 -f 0.0 -p 10 "
        "-R 4")
    assert cmds[1][cmds[1].index("-v") + 1] == "0.0005000000"
    assert cmds[2][3:7] == ["-n", "128", "-n", "128"]


def test_commands_parse_in_the_driver():
    j = launch.read_ini(INI)
    for cmd in launch.solver_commands(j):
        d = host.parse_cli(cmd[1:]).as_dict()
        assert d["u0"] == "TG_VORT" and d["save_every"] == 10 and d["dt"] == 0.001
        assert d["ekmn_drag_flag"] == 1 and d["ekmn_drag_pow"] == 0.0 and d["hyper_visc_pow"] == 2.0
        assert d["trans_iters_flag"] == 0 and d["adapt_step_flag"] == 0
        # run_solver.py:116's default process count, for the initial condition
        assert d["ic_ranks"] == 4


def test_solver_procs_selects_the_rank_count(tmp_path):
    text = open(INI).read().replace("[JOB]\n", "[JOB]\nsolver_procs = 8\n")
    ini = tmp_path / "job.ini"
    ini.write_text(text)
    j = launch.read_ini(str(ini))
    assert j.solver_procs == 8
    for cmd in launch.solver_commands(j):
        assert cmd[-2:] == ["-R", "8"]
        assert host.parse_cli(cmd[1:]).as_dict()["ic_ranks"] == 8


def test_cmdonly_prints(capsys):
    assert launch.main(["-i", INI, "--cmdonly"]) == 0
    out = capsys.readouterr().out
    assert out.count("-i TG_VORT") == 4
