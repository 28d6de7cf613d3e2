"""Write the golden fixtures of tests/golden/ by running the REFERENCE.

Runs oracle/_ref/merlict_golden (the reference's own sources compiled by
oracle/Makefile, with the Philox stream plugged into merlict's
random::Generator) on seeded inputs and stores inputs and outputs as small
.npz files.  Run in the container that has the reference mounted:

    make -C oracle ref && python tests/golden/make_golden.py [fact | plenoscope | zoo | corsika | signal | nsb | calibration | render]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE.parent))

import photon_sets as sources  # noqa: E402
import pyoracle as O  # noqa: E402
from merlict_development_kit_amd import SceneDescription  # noqa: E402

PHOTON_KEYS = O.PHOTON_FIELDS
RESULT_KEYS = ("final_frame", "sensor_id", "final_interaction", "num_interactions", "error", "time_of_flight",
               "x", "y", "cx", "cy", "accumulated_distance")


def save_propagation(name, desc, photons, seed, max_int=5):
    ref = O.ref_propagate(desc, photons, seed=seed, max_num_interactions=max_int)
    arrays = {f"in_{k}": np.asarray(photons[k]) for k in PHOTON_KEYS}
    arrays.update({f"out_{k}": ref[k] for k in RESULT_KEYS})
    arrays["seed"] = np.array([seed], np.uint64)
    arrays["max_num_interactions"] = np.array([max_int], np.uint32)
    np.savez_compressed(HERE / f"{name}.npz", **arrays)
    print(name, "photons", len(photons["support_x"]), "detected", int((ref["sensor_id"] >= 0).sum()))


def probe_rays(n, seed):
    """Rays from above the FACT dish towards random points of the dish and of
    the camera plane."""
    rng = np.random.default_rng(seed)
    tx = rng.uniform(-2.2, 2.2, n)
    ty = rng.uniform(-2.2, 2.2, n)
    tz = np.where(rng.uniform(size=n) < 0.7, 0.1, 4.889)
    sx = rng.uniform(-1.0, 1.0, n)
    sy = rng.uniform(-1.0, 1.0, n)
    sz = np.full(n, 12.0)
    d = np.stack([tx - sx, ty - sy, tz - sz], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return {"support_x": sx, "support_y": sy, "support_z": sz, "direction_x": d[:, 0], "direction_y": d[:, 1],
            "direction_z": d[:, 2], "wavelength": np.full(n, 433e-9)}


def fact_fixtures():
    """The FACT telescope (pixel table and mirror reflectivity of the FACT
    propagation app): propagation with bare pixels and with light guides,
    the clustered tree, probe-ray first hits."""
    fact = SceneDescription.fact(light_guides=False)
    save_propagation("fact_propagation", fact, sources.fact_bunch(2000, seed=3).numpy(), seed=11)
    fact_lg = SceneDescription.fact(light_guides=True)
    save_propagation("fact_lg_propagation", fact_lg, sources.fact_bunch(2000, seed=5).numpy(), seed=13)
    # deeper histories: a higher interaction limit on the light-guide camera
    save_propagation("fact_lg_max10_propagation", fact_lg, sources.fact_bunch(1000, seed=6).numpy(), seed=17,
                     max_int=10)

    tree = O.ref_tree(fact)
    np.savez_compressed(HERE / "fact_tree.npz", **tree)
    print("fact tree", len(tree["frame"]))

    rays = probe_rays(2000, seed=21)
    hits = O.ref_first_hit(fact, rays)
    np.savez_compressed(HERE / "fact_first_hit.npz", **{f"in_{k}": rays[k] for k in PHOTON_KEYS},
                        out_frame=hits["frame"], out_distance=hits["distance"], out_position=hits["position"])
    print("probe rays", len(rays["support_x"]), "hits", int((hits["frame"] >= 0).sum()))


def plenoscope_fixtures():
    """Plenoscope propagation: calibration photons and one bunch with a
    pointing offset, drawn by the library's Philox sources (restated by the
    oracle), up to 10 interactions as plenoscope_propagation.cpp:95."""
    from merlict_development_kit_amd import sources as S
    pleno = SceneDescription.plenoscope()
    cal = O.generate_photons(S.calibration().to_c(O.Source), 2000, seed=1)
    save_propagation("plenoscope_calibration_propagation", pleno, cal, seed=3, max_int=10)
    bunch = O.generate_photons(S.plenoscope_bunch(0.02, -0.035).to_c(O.Source), 2000, seed=2)
    save_propagation("plenoscope_bunch_propagation", pleno, bunch, seed=5, max_int=10)


def zoo_fixtures():
    """Every primitive kind (tests/scene_zoo.py): probe-ray first hits and a
    propagation of the same rays as photons."""
    import scene_zoo
    zoo = scene_zoo.zoo()
    rays = scene_zoo.probe_rays(4000, seed=1)
    hits = O.ref_first_hit(zoo, rays)
    np.savez_compressed(HERE / "zoo_first_hit.npz", **{f"in_{k}": rays[k] for k in PHOTON_KEYS},
                        out_frame=hits["frame"], out_distance=hits["distance"], out_position=hits["position"])
    print("zoo probe rays", len(rays["support_x"]), "hits", int((hits["frame"] >= 0).sum()))
    photons = scene_zoo.probe_rays(3000, seed=2)
    save_propagation("zoo_propagation", zoo, photons, seed=23, max_int=8)


def corsika_fixtures():
    """CORSIKA photon bunches through the reference's EventIoPhotonFactory:
    the first 2000 bunches of the reference's test file telescope.dat
    (weights 0.23-1, negative wavelengths as CORSIKA writes them) and 300
    synthetic bunches with weights over [0, 16] (0, 1, 15.5 and 16 among
    them) and both wavelength signs; plus three bunches with bad weights."""
    tel = O.ref_corsika_extract("/root/reference/merlict_corsika/tests/resources/telescope.dat")[:2000]
    rng = np.random.default_rng(11)
    m = 300
    syn = np.zeros((m, 8), np.float32)
    syn[:, 0] = rng.uniform(-3e4, 3e4, m)
    syn[:, 1] = rng.uniform(-3e4, 3e4, m)
    ang = rng.uniform(0, 2 * np.pi, m)
    sin_t = rng.uniform(0, 0.6, m)
    syn[:, 2] = sin_t * np.cos(ang)
    syn[:, 3] = sin_t * np.sin(ang)
    syn[:, 4] = rng.uniform(-50, 200, m)
    syn[:, 5] = rng.uniform(5e5, 2e6, m)
    syn[:, 6] = rng.uniform(0, 16, m)
    syn[:4, 6] = [0.0, 1.0, 15.5, 16.0]
    syn[:, 7] = rng.uniform(250, 700, m) * np.where(rng.uniform(size=m) < 0.5, -1, 1)
    bunches = np.concatenate([tel, syn])
    ref = O.ref_corsika(bunches, seed=21, id_offset=5)
    assert not ref["status"].any()
    bad = bunches[:3].copy()
    bad[:, 6] = [-0.1, 16.1, 0.5]
    ref_bad = O.ref_corsika(bad, seed=21)
    arrays = {"bunches": bunches, "count": ref["count"], "truth": ref["truth"], "seed": np.array([21], np.uint64),
              "id_offset": np.array([5], np.uint64), "bad_bunches": bad, "bad_status": ref_bad["status"]}
    arrays.update({f"out_{k}": v for k, v in ref["photons"].items()})
    np.savez_compressed(HERE / "corsika_photons.npz", **arrays)
    print("corsika_photons", len(bunches), "bunches", len(ref["truth"]), "photons")


def signal_inputs(n=5000, num_sensors=40, seed=13):
    """Synthetic detected photons for the signal chain: sensor ids with some
    out of range, error flags, distinct arrival times over 0-60 ns,
    wavelengths 280-700 nm, truth ids; a quantum efficiency table over
    200-1200 nm; dark rate, crosstalk, exposure, slices."""
    rng = np.random.default_rng(seed)
    sensor = rng.integers(-3, num_sensors + 3, n).astype(np.int32)
    error = (rng.uniform(size=n) < 0.02).astype(np.uint8)
    t = rng.uniform(0.0, 60e-9, n)
    lam = rng.uniform(280e-9, 700e-9, n)
    truth = rng.integers(0, 1_000_000, n).astype(np.int32)
    qe_x = np.array([1200e-9, 200e-9, 300e-9, 400e-9, 500e-9, 600e-9, 800e-9, 1000e-9])  # unsorted, as merlict allows
    qe_y = np.array([0.02, 0.0, 0.25, 0.4, 0.35, 0.2, 0.08, 0.05])
    return dict(sensor_id=sensor, error=error, time_of_flight=t, wavelength=lam, truth_id=truth,
                num_sensors=np.array([num_sensors], np.int32), qe_x=qe_x, qe_y=qe_y,
                dark_rate=np.array([2e7]), crosstalk=np.array([0.1]), exposure_time=np.array([100e-9]),
                seed=np.array([9], np.uint64), duration=np.array([0.5e-9]), std_dev=np.array([0.4e-9]))


def signal_fixtures():
    """The reference's signal chain (PipelinePhoton sorting, the
    photo-electric converter, pulse extraction) on the synthetic inputs."""
    a = signal_inputs()
    det = (a["error"] == 0)
    ref = O.ref_signal(np.where(det, a["sensor_id"], -1), a["time_of_flight"], a["wavelength"], a["truth_id"],
                       int(a["num_sensors"][0]), a["qe_x"], a["qe_y"], float(a["dark_rate"][0]),
                       float(a["crosstalk"][0]), float(a["exposure_time"][0]), int(a["seed"][0]),
                       float(a["duration"][0]), float(a["std_dev"][0]))
    arrays = {f"in_{k}": v for k, v in a.items()}
    arrays.update({f"out_{k}": v for k, v in ref.items()})
    np.savez_compressed(HERE / "signal_chain.npz", **arrays)
    print("signal_chain", int(ref["pipe_offsets"][-1]), "photons", int(ref["pulse_offsets"][-1]), "pulses",
          int(ref["extracted_offsets"][-1]), "extracted")


NSB_GEOMETRY = (10.0, 3.0, 1.0, 0.2, 3)  # focal length, aperture radius, FoV [deg], pixel FoV [deg], paxels


def nsb_fixtures():
    """The reference's night-sky-background injection on synthetic pipelines
    of the lixels of a small light-field sensor, with the flux of the
    plenoscope propagation example card (apps/examples/PlenoscopePropagation/
    propagation_input_card.json, kept as tests/golden/
    plenoscope_propagation_input_card.json)."""
    card = json.loads((HERE / "plenoscope_propagation_input_card.json").read_text())
    flux = np.array(card["night_sky_background_ligth"]["flux_vs_wavelength"], np.float64)
    exposure = float(card["night_sky_background_ligth"]["exposure_time"])
    meta, _ = O.ref_lightfield(*NSB_GEOMETRY, 1.49, 0.0, 1.1)
    ns = int(meta["num_lixel"])
    rng = np.random.default_rng(23)
    counts = rng.poisson(6.0, ns)
    counts[:3] = 0
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    m = int(off[-1])
    t = np.concatenate([np.sort(rng.normal(30e-9, 3e-9, c)) for c in counts]) if m else np.zeros(0)
    lam = rng.uniform(280e-9, 700e-9, m)
    truth = rng.integers(0, 100000, m).astype(np.int32)
    eff = rng.uniform(0.5, 1.0, ns).astype(np.float32)
    eff[5] = 0.0
    arrays = {"geometry": np.array(NSB_GEOMETRY, np.float64), "flux_x": flux[:, 0].copy(),
              "flux_y": flux[:, 1].copy(), "exposure_time": np.array([exposure]),
              "seed": np.array([17], np.uint64), "in_offsets": off, "in_time": t, "in_wavelength": lam,
              "in_truth": truth, "efficiencies": eff}
    ref = O.ref_nsb(off, t, lam, truth, eff, flux[:, 0], flux[:, 1], exposure, 17, NSB_GEOMETRY)
    arrays.update({f"out_{k}": np.asarray(v) for k, v in ref.items()})
    # no pipeline photon: the exposure starts at -exposure / 2
    z = O.ref_nsb(np.zeros(ns + 1, np.int64), np.zeros(0), np.zeros(0), np.zeros(0, np.int32), eff, flux[:, 0],
                  flux[:, 1], exposure, 18, NSB_GEOMETRY)
    arrays.update({f"empty_{k}": np.asarray(v) for k, v in z.items()})
    np.savez_compressed(HERE / "nsb_injection.npz", **arrays)
    print("nsb_injection", ns, "lixels", m, "photons ->", int(ref["offsets"][-1]), "rate", ref["rate"],
          "start", ref["start"], "| empty ->", int(z["offsets"][-1]), "start", z["start"])


def calibration_records(num_lixel=60, n=1200, seed=29):
    """Synthetic CalibrationPhotonResult records: about 80 % reached, lixel
    ids skewed so that some lixels get none or one record, float values of
    the calibration's scales."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, O.CALIBRATION_PHOTON)
    r["reached_sensor"] = rng.uniform(size=n) < 0.8
    r["lixel_id"] = np.minimum((rng.pareto(1.5, n) * 6).astype(np.int64), num_lixel - 1)
    r["lixel_id"][~r["reached_sensor"].astype(bool)] = rng.integers(0, 1 << 31, (~r["reached_sensor"].astype(bool)).sum())
    r["x_pos_on_principal_aperture"] = rng.uniform(-37, 37, n)
    r["y_pos_on_principal_aperture"] = rng.uniform(-37, 37, n)
    r["x_tilt_vs_optical_axis"] = rng.normal(0, 0.03, n)
    r["y_tilt_vs_optical_axis"] = rng.normal(0, 0.03, n)
    r["time_of_flight"] = rng.normal(5.5e-7, 3e-9, n)
    return r


def calibration_fixtures():
    """The reference's LixelStatisticsFiller (3 blocks of 400 records, 60
    lixels) and its calibration::write bytes."""
    rec = calibration_records()
    raw = O.ref_lixelstats(rec, 60, 3, 400)
    np.savez_compressed(HERE / "lixel_statistics.npz", records=rec.view(np.uint8),
                        config=np.array([60, 3, 400], np.int64), written=np.frombuffer(raw, np.uint8))
    st = np.frombuffer(raw, O.LIXEL_STATISTIC)
    print("lixel_statistics", len(rec), "records", len(st), "lixels", int(np.isnan(st["cx_std"]).sum()), "NaN std")


def render_fixtures():
    """The reference's PinHoleCamera + Tracer (merlict_visual) per pixel."""
    from merlict_development_kit_amd import visual as V
    from render_cases import render_cases, render_colors
    arrays = {}
    for name, desc, cam, gi, seed in render_cases():
        cfg = V.RenderConfig(global_illumination=gi)
        col = render_colors(desc.num_frames)
        arrays[name] = O.ref_render(desc, cam.position, cam.rotation, cam.field_of_view, cam.num_cols,
                                    cam.num_rows, gi, cfg.incoming_direction, cfg.sky_color, col, seed)
        print("render", name, arrays[name].shape, len(np.unique(arrays[name].reshape(-1, 3), axis=0)), "colours")
    np.savez_compressed(HERE / "render.npz", **arrays)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/merlict_golden is not built")
    if len(sys.argv) > 1 and sys.argv[1] == "render":
        render_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "calibration":
        calibration_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "nsb":
        nsb_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "signal":
        signal_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "corsika":
        corsika_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "plenoscope":
        plenoscope_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "zoo":
        zoo_fixtures()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "fact":
        fact_fixtures()
        render_fixtures()
        return
    lens = SceneDescription.lens()
    save_propagation("lens_propagation", lens, sources.lens_beam(2000, seed=1).numpy(), seed=7)
    fact_fixtures()

    fact_facets = O.ref_reflector(4.889, 0.5, 2.1, 0.25, 0, 0.3, 0.01)
    pleno_facets = O.ref_reflector(106.5, 0.0, 35.5, 3.5, 0, 0.75, 0.02)
    hex_facets = O.ref_reflector(4.889, 0.5, 2.1, 0.25, 1, 0.3, 0.01)
    np.savez_compressed(HERE / "reflector_facets.npz", fact=fact_facets, plenoscope=pleno_facets,
                        fact_hex_outline=hex_facets)
    print("facets", len(fact_facets), len(pleno_facets), len(hex_facets))

    lens_cases = np.array([[1.0, 0.25, 1.49], [0.5, 0.1, 1.33], [2.0, 0.3, 1.6]])
    R = np.array([O.ref_lens_maker(*c) for c in lens_cases])
    meta, rows = O.ref_lightfield(106.5, 35.5, 6.5, 0.06667, 9, 1.49, 0.8, 1.1)
    # first rows (lens array head) and summary of the whole light-field sensor
    np.savez_compressed(HERE / "geometry.npz", lens_cases=lens_cases, lens_R=R,
                        lightfield_lens_R=np.array([meta["lens_R"]]), lightfield_lens_f=np.array([meta["lens_f"]]),
                        lightfield_num_lixel=np.array([meta["num_lixel"]]), lightfield_head=rows[:4000],
                        lightfield_num_rows=np.array([len(rows)]),
                        lightfield_pos_sum=rows[:, 3:6].sum(axis=0), lightfield_radius_sum=np.array([rows[:, -1].sum()]))
    print("light field rows", len(rows), meta)
    plenoscope_fixtures()
    zoo_fixtures()


if __name__ == "__main__":
    main()
