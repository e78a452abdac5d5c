"""Video command line (ml-rs/main.rs:170-365) on the CPU: argument handling, creature discovery,
the reference's encoder arguments.  Frame content is checked on the GPU (test_gpu_video.py)."""
import os


def test_refuses_bad_rates(cn):
    assert cn.video._main(["--creature", "x", "--fps", "0"]) == 2
    assert cn.video._main(["--with-sound", "--creature", "x", "--seconds", "0"]) == 2


def test_mux_arguments_match_the_reference(cn):
    assert cn.video.mux_args("v/x.mp4", "v/x.wav", "v/x_muxed.mp4") == [
        "ffmpeg", "-y", "-i", "v/x.mp4", "-i", "v/x.wav", "-c:v", "copy", "-c:a", "aac", "-shortest", "v/x_muxed.mp4"]


def test_no_creatures_found(cn, tmp_path, monkeypatch, capsys):
    monkeypatch.chdir(tmp_path)
    assert cn.video._main([]) == 1
    assert "No creatures found in seed/" in capsys.readouterr().err


def test_discover_creatures_sorted_json_only(cn, tmp_path):
    os.makedirs(tmp_path / "seed")
    for f in ("b.json", "a.json", "notes.txt", "c.json.bak"):
        (tmp_path / "seed" / f).write_text("{}")
    assert cn.video.discover_creatures(str(tmp_path)) == ["a", "b"]
    assert cn.video.discover_creatures(str(tmp_path / "nowhere")) == []


def test_encoder_arguments_match_the_reference(cn):
    assert cn.video.ffmpeg_args(512, 30, "videos/x.mp4") == [
        "ffmpeg", "-y", "-f", "rawvideo", "-pixel_format", "rgb24", "-video_size", "512x512", "-framerate", "30",
        "-i", "-", "-c:v", "libx264", "-crf", "18", "-preset", "medium", "-pix_fmt", "yuv420p", "videos/x.mp4"]


def test_existing_video_is_skipped(cn, tmp_path):
    os.makedirs(tmp_path / "videos")
    (tmp_path / "videos" / "x.mp4").write_bytes(b"")
    logs = []
    assert cn.video.render_video("x", output=str(tmp_path / "videos"), log=logs.append) is None
    assert logs == ["  Video already exists, skipping."]
