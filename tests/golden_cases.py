"""The seeded scenes whose oracle outputs are hashed in
tests/golden/oracle_hashes.json."""
from polyred_amd import scenes

from . import cases as edge


def c1_shadow_ao():
    s = scenes.c1_plain()
    s.name = "c1_shadow_ao"
    s.lights[0].cast_shadow = True
    s.shadow_map = True
    for m in s.materials:
        m.receive_shadow = True
        m.ambient_occlusion = True
    return s


def cases_dict():
    d = {"c1_plain": scenes.c1_plain(), "c1_shadow_ao": c1_shadow_ao(),
         "big_triangle": edge.big_triangle(), "depth_tie": edge.depth_tie(),
         "nan_first": edge.nan_depth(True), "nan_second": edge.nan_depth(False),
         "no_lights": edge.no_lights(), "no_materials": edge.no_materials()}
    for seed in range(4):
        d[f"tiny_{seed}"] = scenes.tiny(seed, width=40, height=30, n_tris=12)
    for seed in range(2):
        d[f"tiny_ortho_{seed}"] = scenes.tiny(100 + seed, width=40, height=30, n_tris=12, perspective=False)
    d["textured"] = scenes.textured(seed=6, width=96, height=72)
    d["textured_ortho_nearest"] = scenes.textured(seed=7, width=96, height=72, perspective=False, use_mipmap=False)
    # render/raster_test.go's TestRender scene: the textured bunny, 1920x1080, MSAA 2
    d["reference_bunny"] = edge.reference_bunny()
    # the scenes of the reference's CPU-vs-GPU deferred tests (render/gpudeferred_*_test.go)
    d["reference_shadow"] = edge.reference_shadow()
    d["reference_shadow2"] = edge.reference_shadow(two_lights=True)
    d["reference_multimaterial"] = edge.reference_multimaterial()
    d["reference_ao"] = edge.reference_ao()
    d["reference_directional"] = edge.reference_directional()
    # the scenes whose shadows and AO the oracle tests pin in closed form
    # (tests/test_oracle_shadow.py, tests/test_oracle_ao.py)
    d["ao_step_x"] = edge.ao_step("x")
    d["ao_step_y"] = edge.ao_step("y")
    d["shadow_footprint"] = edge.shadow_footprint()
    d["shadow_footprint_nc"] = edge.shadow_footprint(non_casting=True)
    d["shadow_footprint_bf"] = edge.shadow_footprint(back_face=True)
    return d


def all_cases():
    return cases_dict()
