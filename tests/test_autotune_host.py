"""The march autotuner's candidate set and choice, on a stub context (no GPU): the
binary16 texture layout is a candidate only where the sources stay put between frames,
prefetch only where a prefetch kernel exists, and the fastest candidate is kept."""
import numpy as np

from isaac_amd import autotune


class _Stub:
    TUNE_MARCH_PREFETCH, TUNE_MARCH_WAVES, TUNE_MARCH_TILES, TUNE_TEX_LAYOUT, TUNE_MARCH_THREADS = 0, 1, 2, 3, 4
    FORM_COMBINED = 0

    def __init__(self, best):
        self.render_form = 0
        self.tex_layout = 0
        self.framebuffer_size = (64, 32)
        self.local_size = (16, 16, 16)
        self.volume_sources, self.particle_sources = [1], []
        self.interpolation = True
        self.state = {}
        self.tried = []
        self.best = best

    def set_tuning(self, key, value):
        self.state[key] = value
        if key == self.TUNE_TEX_LAYOUT:
            self.tex_layout = value

    def update_buffers(self):
        pass

    def _update_modelview(self):
        pass

    def draw(self, out=None):
        s = self.state
        self.tried.append((s.get(3, 0), s.get(4, 256), s.get(1, 0), s.get(0, 0), s.get(2, 4)))

    def timings(self):
        s = self.state
        cand = (s.get(3, 0), s.get(4, 256), s.get(1, 0), s.get(0, 0), s.get(2, 4))
        return {"march_us": 100.0 if cand == self.best else 200.0, "optimization_buffer_us": 10.0}


class _Pose:
    look_at = (0.0, 0.0, 0.0)
    rotation = np.eye(3)
    distance = -4.0


def test_layout_16_only_for_static_sources():
    for update_each_frame in (True, False):
        autotune._CACHE.clear()
        vis = _Stub(best=(0, 64, 6, 0, 4))
        autotune.tune_march(vis, [_Pose()], frames=1, use_cache=False, update_each_frame=update_each_frame)
        layouts = {t[0] for t in vis.tried}
        assert (16 in layouts) == (not update_each_frame), layouts
        # prefetch only at the compiler's default occupancy and never with the brick layout
        assert all(t[3] == 0 or (t[2] == 0 and t[0] != 8) for t in vis.tried)


def test_the_fastest_candidate_is_kept():
    autotune._CACHE.clear()
    vis = _Stub(best=(16, 64, 0, 1, 8))
    r = autotune.tune_march(vis, [_Pose()], frames=2, use_cache=False, update_each_frame=False)
    assert r["choice"] == {"march_waves": 0, "march_prefetch": 1, "march_tiles": 8, "tex_layout": 16, "march_threads": 64}
    assert vis.tex_layout == 16
