"""CPU: the tuning table (set / get / list / clear) and its JSON cache."""
import json

from dosecuda_amd import autotune


def test_table_and_cache_roundtrip(tmp_path):
    autotune.clear()
    autotune.set_choice("raytrace", "100,100,100,1", 3, 0.25)
    autotune.set_choice("conv3d", "1,64,64,64,32,0,0,16,3", 1001, 0.04)
    autotune.set_choice("raytrace", "100,100,100,1", 5, 0.2)      # replaces
    assert autotune.get("raytrace", "100,100,100,1") == 5
    assert autotune.get("pencil", "1,1,1,0") is None
    e = autotune.entries()
    assert len(e) == 2
    p = str(tmp_path / "tune.json")
    autotune.save(p)
    doc = json.load(open(p))
    assert doc["arch"] == autotune.arch() and len(doc["entries"]) == 2
    autotune.clear()
    assert autotune.entries() == []
    assert autotune.load(p) == 2
    assert autotune.get("conv3d", "1,64,64,64,32,0,0,16,3") == 1001
    autotune.clear()


def test_cache_of_another_arch_is_ignored(tmp_path):
    autotune.clear()
    p = tmp_path / "t.json"
    p.write_text(json.dumps(dict(arch="gfx000", entries=[dict(kernel="raytrace", key="1,1,1,0", choice=2)])))
    assert autotune.load(str(p)) == 0
    assert autotune.entries() == []


def test_surrogate_layer_shapes():
    from dosecuda_amd.surrogate import SurrogateNet, IN_CH
    shapes = autotune.surrogate_layer_shapes(SurrogateNet(), (64, 64, 64))
    assert len(shapes) == 15
    assert shapes[0] == (1, 64, 64, 64, IN_CH, 0, 0, 16, 3)
    assert shapes[8] == (1, 16, 16, 16, 128, 64, 1, 64, 3)      # d2.0: upsampled bottleneck + skip
    assert shapes[-1] == (1, 64, 64, 64, 16, 0, 0, 16, 1)
