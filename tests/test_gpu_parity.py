"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on
identical seeded inputs. The bar is exact agreement of every buffer (float
values; raw 16-bit codes for FP16S/FP16C DDFs), step by step."""
import numpy as np
import pytest

from harness import compare_all, make_pair

pytestmark = pytest.mark.gpu


def run_both(gpu, orc, steps):
    gpu.initialize()
    orc.initialize()
    for _ in range(steps):
        gpu.do_time_step()
        orc.do_time_step()


@pytest.mark.parametrize("float_type", [2, 0, 1])
@pytest.mark.parametrize("trt", [0, 1])
def test_fluid_d3q19(float_type, trt):
    rng = np.random.default_rng(100 + float_type * 10 + trt)
    gpu, orc, info = make_pair(rng, nx=20, ny=12, nz=9, float_type=float_type, relaxation_time=trt, ffield=True)
    run_both(gpu, orc, 5)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("vs,nz", [(0, 1), (0, 3), (1, 8)])
def test_fluid_other_sets(vs, nz):
    rng = np.random.default_rng(200 + vs)
    gpu, orc, info = make_pair(rng, nx=16, ny=10, nz=nz, velocity_set=vs, float_type=2)
    run_both(gpu, orc, 4)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("float_type", [2, 0, 1])
@pytest.mark.parametrize("lod_depth", [1, 2, 3])
def test_mhd_d3q19(float_type, lod_depth):
    rng = np.random.default_rng(300 + float_type * 10 + lod_depth)
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=16, mhd=True, lod_depth=lod_depth, float_type=float_type)
    run_both(gpu, orc, 3)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("float_type,trt", [(2, 0), (1, 1), (0, 0)])
def test_mhd_electron_density0(float_type, trt):
    """The electron gas started at a non-zero density (build extension
    electron_density0), every buffer exact."""
    rng = np.random.default_rng(350 + float_type * 10 + trt)
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=16, mhd=True, lod_depth=2, float_type=float_type,
                               relaxation_time=trt, electron_density0=0.75)
    run_both(gpu, orc, 4)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("nx,ny,nz,D,ft", [(512, 4, 4, 2, 1), (384, 8, 8, 2, 2), (1024, 2, 4, 1, 0)])
def test_mhd_rows_wider_than_a_block(nx, ny, nz, D, ft):
    """Rows longer than 256 cells: stream_collide blocks hold part of a row
    (whole LOD segments), every buffer exact."""
    rng = np.random.default_rng(500 + nx + D)
    gpu, orc, info = make_pair(rng, nx=nx, ny=ny, nz=nz, mhd=True, lod_depth=D, float_type=ft)
    run_both(gpu, orc, 3)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("nx,ny,nz,D,ft", [(24, 12, 12, 2, 2), (24, 12, 12, 2, 1), (24, 12, 12, 2, 0),
                                           (40, 20, 10, 1, 1), (96, 6, 6, 1, 2), (300, 4, 4, 2, 1)])
def test_mhd_rows_and_lod_blocks_not_powers_of_two(nx, ny, nz, D, ft):
    """Row lengths and LOD block sides that are not powers of two (as the
    384^3 workload's 24-cell blocks): stream_collide blocks of 216, 200, 192
    and 150 work-items (the last wavefront partial), the periodic wrap of the
    +x hand-over inside whole rows, and LOD segments of 6, 20, 48 and 75
    cells summed by the pairwise tree. Every buffer exact."""
    rng = np.random.default_rng(700 + nx + ny + D + ft)
    gpu, orc, info = make_pair(rng, nx=nx, ny=ny, nz=nz, mhd=True, lod_depth=D, float_type=ft)
    run_both(gpu, orc, 3)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("ft", [2, 1])
def test_mhd_lod_depth5_near_field_spans_the_lattice(ft):
    """LOD depth 5 at 32^3: single-cell LOD blocks (8^5 top-level sources,
    one-cell LOD segments in stream_collide) and a near-field block side of
    n_x / 2^(2^5) cells, whose shift OpenCL takes modulo 32 (the block
    side of sim_kernels.cl:907-938), so the near field spans the whole
    lattice. Every buffer exact after 2 steps."""
    rng = np.random.default_rng(800 + ft)
    gpu, orc, info = make_pair(rng, nx=32, ny=32, nz=32, mhd=True, lod_depth=5, float_type=ft)
    run_both(gpu, orc, 2)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("shape,D", [((64, 64, 64), 3), ((128, 64, 64), 3), ((192, 64, 64), 3), ((32, 64, 64), 3),
                                     ((64, 64, 32), 2), ((96, 64, 64), 2)])
def test_exact_field_on_lod_tiles(shape, D):
    """EXACT field arithmetic, every buffer exact against the oracle. At D = 3
    the near field is off (its block side n / 2^(2^D) is below one cell) and
    the LOD blocks (8, 16 and 24 cells along x) take the tiled kernel with 8
    cells per work-item (tiles of 1 x 8 x 8, 2 x 8 x 8 and 3 x 8 x 8
    work-items), 4-cell blocks the one with 4 cells per work-item (1 x 8 x 8);
    at D = 2 these lattices have a near field, which the generic kernel
    evaluates."""
    rng = np.random.default_rng(600 + shape[0])
    gpu, orc, info = make_pair(rng, nx=shape[0], ny=shape[1], nz=shape[2], mhd=True, lod_depth=D, float_type=2)
    run_both(gpu, orc, 2)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("float_type,trt,D,inlet,graphics", [(2, 0, 2, False, True), (1, 0, 2, False, True),
                                                              (0, 1, 1, False, True), (2, 0, 0, False, True),
                                                              (1, 0, 2, True, True), (2, 1, 1, False, False)])
def test_mhd_subgrid_ecr(float_type, trt, D, inlet, graphics):
    """SUBGRID_ECR (electron temperature, ECR heating by E_var, gyration
    drift, ionisation) with TYPE_E cells, with the charge inlet, and with
    field updates off (Et then not written), every buffer exact."""
    rng = np.random.default_rng(700 + 10 * float_type + trt + D + 100 * inlet + 200 * (not graphics))
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=16, mhd=True, lod_depth=D, float_type=float_type,
                               relaxation_time=trt, ecr=True, charge_inlet=inlet, graphics=graphics)
    run_both(gpu, orc, 3)
    assert compare_all(gpu, orc, info) == []


def test_mhd_lod_depth0_and_d3q15():
    rng = np.random.default_rng(400)
    gpu, orc, info = make_pair(rng, nx=8, ny=8, nz=8, mhd=True, lod_depth=0, float_type=2)
    run_both(gpu, orc, 2)
    assert compare_all(gpu, orc, info) == []
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=8, mhd=True, lod_depth=2, float_type=2, velocity_set=1)
    run_both(gpu, orc, 2)
    assert compare_all(gpu, orc, info) == []


@pytest.mark.parametrize("float_type,trt", [(2, 0), (1, 0), (2, 1)])
def test_mhd_charge_inlet(float_type, trt):
    """Charge inlet extension: TYPE_E cells hold their charge and reset
    their charge DDFs to equilibrium, every buffer exact."""
    rng = np.random.default_rng(450 + float_type + 5 * trt)
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=16, mhd=True, lod_depth=2, float_type=float_type,
                               relaxation_time=trt, charge_inlet=True)
    run_both(gpu, orc, 4)
    assert compare_all(gpu, orc, info) == []


def test_device_codecs_match_oracle():
    import oracle_lib as O
    import ionsolver_amd as I
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(1 << 20) * np.exp2(rng.integers(-40, 3, 1 << 20))).astype(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, 1.99951171875, 2.0, -2.0, 6.1035156e-05, 2.9802322e-08, 1e-30,
                        65504.0 / 32768, 65519.0 / 32768, 65520.0 / 32768, 3.0, np.inf, -np.inf], np.float32)
    x = np.concatenate([x, special])
    codes = np.arange(65536, dtype=np.uint16)
    x = np.concatenate([x, np.array([np.nan, -np.nan], np.float32),
                        np.array([0x7FC00001, 0xFFC12345, 0x7F800001], np.uint32).view(np.float32)])
    for ft, enc in ((1, O.fp16c_encode), (0, O.fp16s_store)):
        g, o = I.device_codec(ft, 0, x), enc(x)
        bad = np.flatnonzero(g != o)
        assert bad.size == 0, (ft, bad.size, x[bad[:8]], g[bad[:8]], o[bad[:8]])
    # FP16C encoder around its normal / denormal / flush boundaries: every
    # mantissa for the exponent fields 100, 101, 111, 112 and 113 (the +0x800
    # rounding carry moves 112 into 113), a strided sweep over 96..130
    m = np.arange(1 << 23, dtype=np.uint32)
    full = np.concatenate([(e << 23) | m for e in (100, 101, 111, 112, 113)])
    strided = np.concatenate([(e << 23) | m[::13] for e in range(96, 131)])
    sweep = np.concatenate([full, strided])
    sweep = np.concatenate([sweep, sweep | np.uint32(0x80000000)]).view(np.float32)
    g, o = I.device_codec(1, 0, sweep), O.fp16c_encode(sweep)
    bad = np.flatnonzero(g != o)
    assert bad.size == 0, (bad.size, sweep[bad[:8]].view(np.uint32), g[bad[:8]], o[bad[:8]])
    # the denormal codes' rounding steps: the encoder rounds |x| 2^25 half up, so
    # the code changes at x = (k + 1/2) 2^-25; the float below, at and above every
    # such point, k = 0 .. 2047, both signs
    k = np.arange(2048, dtype=np.float64)
    mid = ((k + 0.5) * 2.0 ** -25).astype(np.float32)
    steps = np.concatenate([np.nextafter(mid, np.float32(0)), mid, np.nextafter(mid, np.float32(1))])
    steps = np.concatenate([steps, -steps])
    g, o = I.device_codec(1, 0, steps), O.fp16c_encode(steps)
    bad = np.flatnonzero(g != o)
    assert bad.size == 0, (bad.size, steps[bad[:8]], g[bad[:8]], o[bad[:8]])
    for ft, dec in ((1, O.fp16c_decode), (0, O.fp16s_load)):
        g, o = I.device_codec(ft, 1, codes), dec(codes)
        assert np.array_equal(g.view(np.uint32)[~np.isnan(o)], o.view(np.uint32)[~np.isnan(o)])
        assert np.array_equal(np.isnan(g), np.isnan(o))


def test_conservation_sums():
    rng = np.random.default_rng(6)
    gpu, orc, info = make_pair(rng, nx=32, ny=16, nz=8, mhd=True, lod_depth=1, float_type=2)
    m, q = gpu.conservation()
    flags = info["arrays"]["flags"]
    fl = (flags & 0x1F) != 1
    assert abs(m - np.sum(info["arrays"]["rho"][fl], dtype=np.float64)) < 1e-9 * abs(m)
    assert abs(q - np.sum(info["arrays"]["q"][fl], dtype=np.float64)) < 1e-9 * max(1.0, abs(q))


def test_fast_field_mode_tolerance():
    """FAST field arithmetic (hardware rsqrt, fused accumulation) stays within
    rel-L2 1e-5 of the exact field for E - E_stat and B - B_stat."""
    from harness import rel_l2
    from ionsolver_amd import Field, FieldMode
    rng = np.random.default_rng(8)
    gpu, orc, info = make_pair(rng, nx=32, ny=32, nz=32, mhd=True, lod_depth=3, float_type=1)
    gpu.initialize()
    gpu.do_time_step()
    e_stat, b_stat = gpu.read_field(Field.E_STAT), gpu.read_field(Field.B_STAT)
    ex, bx = gpu.read_field(Field.E_DYN) - e_stat, gpu.read_field(Field.B_DYN) - b_stat
    # recompute the same step's field with FAST arithmetic from the same LOD sums
    gpu.set_field_mode(FieldMode.FAST)
    qu = gpu.read_field(Field.QU_LOD)
    gpu2, _, _ = make_pair(np.random.default_rng(8), nx=32, ny=32, nz=32, mhd=True, lod_depth=3, float_type=1)
    gpu2.set_field_mode(FieldMode.FAST)
    gpu2.initialize()
    gpu2.do_time_step()
    assert np.array_equal(gpu2.read_field(Field.QU_LOD), qu)
    ef = gpu2.read_field(Field.E_DYN) - e_stat
    bf = gpu2.read_field(Field.B_DYN) - b_stat
    assert rel_l2(ef, ex) < 1e-5
    assert rel_l2(bf, bx) < 1e-5


def test_save_load_roundtrip(tmp_path):
    import ionsolver_amd as I
    from ionsolver_amd import Field
    rng = np.random.default_rng(9)
    gpu, orc, info = make_pair(rng, nx=16, ny=8, nz=8, mhd=True, lod_depth=1, float_type=1)
    gpu.initialize()
    gpu.run(2)
    p = str(tmp_path / "state.ion")
    gpu.save(p)
    back = I.Lbm.load(p)
    for f in (Field.FLAGS, Field.RHO, Field.U, Field.Q):
        assert np.array_equal(back.read_field(f), gpu.read_field(f))
    assert int(back.config.float_type) == 1 and back.config.n_x == 16 and back.config.ext_magneto_hydro


@pytest.mark.parametrize("shape,D", [((64, 64, 64), 3), ((128, 64, 64), 3), ((192, 64, 64), 3), ((32, 64, 64), 3),
                                     ((64, 64, 32), 2), ((32, 32, 32), 3)])
def test_fast_field_vs_exact_one_step(shape, D):
    """FAST vs EXACT field arithmetic from identical LOD sums (the LOD-tiled
    kernel serves the D = 3 shapes of 64 cells and more per axis, the generic
    kernel the others, with a near field at D = 2): rel-L2 <= 1e-5 on
    E - E_stat and B - B_stat."""
    from harness import rel_l2
    from ionsolver_amd import Field, FieldMode
    out = {}
    for mode in (FieldMode.EXACT, FieldMode.FAST):
        gpu, _, _ = make_pair(np.random.default_rng(21), nx=shape[0], ny=shape[1], nz=shape[2], mhd=True,
                              lod_depth=D, float_type=2)
        gpu.set_field_mode(mode)
        gpu.initialize()
        gpu.do_time_step()
        out[mode] = [gpu.read_field(f) for f in (Field.QU_LOD, Field.E_DYN, Field.B_DYN, Field.E_STAT, Field.B_STAT)]
    ex, fa = out[FieldMode.EXACT], out[FieldMode.FAST]
    assert np.array_equal(ex[0], fa[0])
    assert rel_l2(fa[1] - fa[3], ex[1] - ex[3]) < 1e-5
    assert rel_l2(fa[2] - fa[4], ex[2] - ex[4]) < 1e-5


@pytest.mark.parametrize("mhd,lod,trt", [(False, 0, 0), (False, 0, 1), (True, 1, 0), (True, 2, 1)])
def test_fp64_instantiation_matches_the_fp64_oracle(mhd, lod, trt):
    """The FP64 instantiation of the kernels (double DDFs, fields and
    arithmetic) against the oracle's float64 build, every buffer exact."""
    rng = np.random.default_rng(300 + 10 * lod + trt)
    gpu, orc, info = make_pair(rng, nx=16, ny=16, nz=8, float_type=3, mhd=mhd, lod_depth=lod, relaxation_time=trt)
    run_both(gpu, orc, 4)
    assert compare_all(gpu, orc, info) == []


def test_storage_format_errors_against_fp64():
    """Taylor-Green vortex, 64^3, 200 steps: rel-L2 of rho - 1 and u of each
    storage format against the FP64 instantiation on the same GPU."""
    from harness import rel_l2
    from ionsolver_amd import Field, FloatType
    from ionsolver_amd.scenarios import apply, fluid_tgv
    out = {}
    for ft in (FloatType.FP64, FloatType.FP32, FloatType.FP16S, FloatType.FP16C):
        cfg, arrays = fluid_tgv(n=64, float_type=ft)
        import ionsolver_amd as I
        lbm = I.Lbm(cfg)
        apply(lbm, arrays)
        lbm.initialize()
        lbm.run(200)
        out[ft] = (lbm.read_field(Field.RHO).astype(np.float64) - 1.0, lbm.read_field(Field.U).astype(np.float64))
    err = {ft: (rel_l2(out[ft][0], out[FloatType.FP64][0]), rel_l2(out[ft][1], out[FloatType.FP64][1]))
           for ft in out if ft != FloatType.FP64}
    print("rel-L2 vs FP64 (rho-1, u):", {k.name: v for k, v in err.items()})
    assert err[FloatType.FP32][1] < 1e-5 and err[FloatType.FP32][0] < 1e-3
    assert err[FloatType.FP16S][1] < 1e-2 and err[FloatType.FP16C][1] < 1e-2
