"""Restart files in the reference's .ers format (Ewoms::Restart, io/restart.hh:36-271;
Simulator::serialize, utils/simulator.hh:923-958; EcfvDiscretization::serialize,
ecfvdiscretization.hh:186-202).

The expected layout is rebuilt here from those rules (magic cookie with the YaspGrid entity
counts, a Simulator section with the clock, one "p_w S_w " line per element, empty section
ends); a run resumed from a restart file continues bit for bit like the uninterrupted run;
a file written the way the reference writes it loads; a file of another grid is refused."""
import numpy as np
import pytest


def _entity_counts(P):
    nx, ny, nz = P.nx, P.ny, P.nz
    if P.dim == 2:
        return nx * ny, nx * (ny + 1) + (nx + 1) * ny, (nx + 1) * (ny + 1)
    return (nx * ny * nz,
            nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz,
            (nx + 1) * (ny + 1) * (nz + 1))


def _reference_text(P, x, clock):
    """The file Restart::serializeBegin + Simulator::serialize + serializeEntities<0>
    write (stream precision 20: 17 significant digits or more round-trip a double)."""
    ne, ned, nv = _entity_counts(P)
    lines = [f"eWoms restart file: gridName='blubb' numCPUs=1 myRank=0 numElements={ne} "
             f"numEdges={ned} numVertices={nv}", "", "Simulator",
             " ".join([str(clock["episode_idx"])] +
                      [f"{clock[k]:.20g}" for k in ("episode_start_time", "episode_length",
                                                    "start_time", "time")] +
                      [str(clock["time_step_idx"])]) + " ",
             "Entities: Codim 0"]
    lines += [f"{p:.20g} {s:.20g} " for p, s in x.reshape(-1, 2)]
    return "\n".join(lines) + "\n\n"


def test_restart_file_name():
    """Restart::restartFileName_ (restart.hh:70-85): the time in the stream's default
    format, '.' as the output directory means none."""
    import ewg
    assert ewg.restart_file_name(".", "lens_immiscible_ecfv_ad", 3000.0) == \
        "lens_immiscible_ecfv_ad_time=3000_rank=0.ers"
    assert ewg.restart_file_name("out", "x", 1234.5678, 3) == "out/x_time=1234.57_rank=3.ers"
    assert ewg.restart_file_name("out/", "x", 2.5e-7) == "out/x_time=2.5e-07_rank=0.ers"


CLOCK = dict(episode_idx=0, episode_start_time=0.0, episode_length=1e100, start_time=0.0,
             time=1500.0, time_step_idx=6)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["lens", "het3d"])
def test_written_file_has_the_reference_layout(tmp_path, which):
    import ewg
    from ewg import problems
    P = (problems.lens_problem(nx=12, ny=8) if which == "lens"
         else problems.heterogeneous_3d(9, 7, 5, seed=3, patch=2))
    g = ewg.GpuContext(P)
    for _ in range(2):
        assert g.newton(P.dt)[0] == 0
        g.advance_time_level()
    path = tmp_path / ewg.restart_file_name(".", "case", 1500.0)
    g.write_restart(path, **CLOCK)
    text = path.read_text()
    x = g.solution(0)
    lines = text.split("\n")
    ref = _reference_text(P, x, CLOCK).split("\n")
    assert lines[:5] == ref[:5]
    assert len(lines) == len(ref) and lines[-2:] == ["", ""]
    vals = np.array([[float(v) for v in ln.split()] for ln in lines[5:5 + P.N]])
    assert np.array_equal(vals.ravel(), x)          # exact round trip of every double
    assert all(ln.endswith(" ") for ln in lines[5:5 + P.N])


@pytest.mark.gpu
def test_resume_from_restart_is_bit_identical(tmp_path):
    """12 steps straight against 6 steps, a restart file, a fresh context loading it,
    6 more steps: the same Newton iterations and the same state bit for bit (the file's
    20-digit decimals round-trip every double; loading sets solution(1) = solution(0))."""
    import ewg
    from ewg import problems
    P = problems.lens_problem()
    a = ewg.GpuContext(P)
    its_a = []
    for _ in range(12):
        rc, it, _ = a.newton(P.dt)
        assert rc == 0
        its_a.append(it)
        a.advance_time_level()
    b = ewg.GpuContext(P)
    its_b = []
    for _ in range(6):
        rc, it, _ = b.newton(P.dt)
        its_b.append(it)
        b.advance_time_level()
    path = tmp_path / "lens_time=1500_rank=0.ers"
    b.write_restart(path, **CLOCK)
    c = ewg.GpuContext(P)
    clock = c.read_restart(path)
    assert clock == CLOCK
    assert np.array_equal(c.solution(0), b.solution(0))
    assert np.array_equal(c.solution(1), b.solution(0))
    for _ in range(6):
        rc, it, _ = c.newton(P.dt)
        assert rc == 0
        its_b.append(it)
        c.advance_time_level()
    print("newton iterations", its_a, its_b, flush=True)
    assert its_a == its_b
    assert np.array_equal(c.solution(0), a.solution(0))


@pytest.mark.gpu
def test_reads_a_file_written_like_the_reference(tmp_path):
    import ewg
    from ewg import problems
    P = problems.heterogeneous_3d(10, 8, 6, seed=4, patch=2)
    x = problems.perturbed_state(P, seed=9).ravel()
    path = tmp_path / "het_time=1500_rank=0.ers"
    path.write_text(_reference_text(P, x, CLOCK))
    g = ewg.GpuContext(P)
    assert g.read_restart(path) == CLOCK
    assert np.array_equal(g.solution(0), x) and np.array_equal(g.solution(1), x)


@pytest.mark.gpu
def test_refuses_a_file_of_another_grid(tmp_path):
    import ewg
    from ewg import problems
    from ewg.capi import SolverError
    P, Q = problems.lens_problem(nx=12, ny=8), problems.lens_problem(nx=16, ny=8)
    path = tmp_path / "lens_time=0_rank=0.ers"
    path.write_text(_reference_text(Q, Q.x0.ravel(), CLOCK))
    g = ewg.GpuContext(P)
    with pytest.raises(SolverError, match="another grid"):
        g.read_restart(path)
    assert np.array_equal(g.solution(0), P.x0.ravel())
