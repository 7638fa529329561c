"""The byte accounting the benchmark reports (alg_bytes_per_site, b_alg),
derived here from the sizes of the buffers each kernel class reads and
writes, at 256^3 FP32 with two fields and the tensor modes."""
import bench

N, S = 256, 4                      # lattice side, bytes per stored real (FP32)
SITES = N ** 3
CPLX = 2 * S                       # bytes of one complex value
HALF = N * N * (N // 2 + 1)        # complex values of one half-spectrum (also the modes)
ROWS = N * N * (N // 2)            # complex slots of one packed row block (kz 0 and N/2 share one)
STATE = 4 * SITES * S              # field and velocity of two fields
NSC = 10                           # stress components stored: (xx - zz, xy, xz, yy - zz, yz) per RK2 stage


def per_site(nbytes):
    return nbytes / SITES


def test_class_bytes_follow_the_buffers():
    got = bench.alg_bytes_per_site(S, N)
    expect = {
        # reads the state, writes the NSC packed row blocks
        "sweep_a": per_site(STATE + NSC * ROWS * CPLX),
        "finalize": 0.0,
        # reads the packed rows, writes them back with the kz = N/2 plane split off
        "stress_fft": per_site(NSC * ROWS * CPLX + NSC * HALF * CPLX),
        # reads the NSC half-spectra; reads and writes h and l, two polarisations each
        "gw_modes": per_site(NSC * HALF * CPLX + 2 * 2 * (2 * HALF * CPLX)),
        # reads the state, writes the new state
        "sweep_b": per_site(2 * STATE),
    }
    assert set(got) == set(expect)
    for k, v in expect.items():
        assert abs(got[k] - v) < 1e-12, (k, got[k], v)
    assert abs(sum(got.values()) - 240.875) < 1e-12


def test_library_rows_add_the_real_stress_round_trip():
    """Without the rows in sweep A, sweep A writes the NSC real stress arrays
    and the transform class reads them, writes the half-spectra and transforms
    them along j in place."""
    got = bench.alg_bytes_per_site(S, N, rows_in_sweep=False)
    real = NSC * SITES * S
    assert abs(got["sweep_a"] - per_site(STATE + real)) < 1e-12
    assert abs(got["stress_fft"] - per_site(real + NSC * HALF * CPLX + 2 * NSC * HALF * CPLX)) < 1e-12


def test_minimal_dataflow_bytes():
    """B_alg, the minimal dataflow: the sweeps read the state twice and write it
    once; the 12 real stress arrays are written and read; one ideal transform
    writes the 12 half-spectra once; the mode update reads them and reads and
    writes the six complex components of h and l."""
    sweeps = 3 * STATE
    stress = 2 * 12 * SITES * S
    transform = 12 * HALF * CPLX
    modes = 12 * HALF * CPLX + 2 * (2 * 6 * HALF * CPLX)
    assert abs(bench.b_alg(S, N) - per_site(sweeps + stress + transform + modes)) < 1e-9
    assert abs(bench.b_alg(S, N) - 337.5) < 1e-12
