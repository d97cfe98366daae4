"""Writes the ruins fixture: the reference's demo scene and its camera
(scenes/ruins/ruins.obj, scenes/ruins/camera.cfg), copied as data; the OBJ is
gzip-compressed (mtime 0, so the bytes are reproducible).  Run in the build
container, where /root/reference exists."""
import gzip
import os
import shutil

SRC = "/root/reference/scenes/ruins"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(SRC, "ruins.obj"), "rb") as f, \
        open(os.path.join(HERE, "ruins.obj.gz"), "wb") as raw, \
        gzip.GzipFile(filename="ruins.obj", mode="wb", fileobj=raw, mtime=0, compresslevel=9) as out:
    shutil.copyfileobj(f, out)
shutil.copyfile(os.path.join(SRC, "camera.cfg"), os.path.join(HERE, "ruins_camera.cfg"))
