"""Writes tests/golden/room_subset/: a shippable part of the reference's
living room (scenes/room.json) -- its camera and point light
(room.json:2656-2672), the named BSDFs and meshes of the shapes listed below
(floor, walls, the four lamp quads whose emission lives in living_room.mtl,
pictures, carpet, books, the TV), the JPEG textures they name and
living_room.mtl -- copied byte for byte; scene.json keeps only those
shapes and BSDFs. Run it where the reference is present:

    python tests/golden/make_room_subset.py /root/reference/scenes
"""
import json
import os
import shutil
import sys

KEEP = ["Floor", "Walls", "MiddleLight", "LeftLight", "RightLight", "BackLight", "Picture1", "Picture2",
        "Picture4", "Picture5", "LargePicture", "Carpet", "Books", "TvScreen", "TvBevel", "BlackMarble"]


def main(src: str) -> None:
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "room_subset")
    os.makedirs(os.path.join(out, "textures"), exist_ok=True)
    with open(os.path.join(src, "room.json")) as f:
        room = json.load(f)
    shapes = [s for s in room["shapes"] if s["Mesh"][1]["Named"] in KEEP]
    bsdfs = {k: v for k, v in room["named_bsdfs"].items() if k in KEEP}
    for s in shapes:
        shutil.copyfile(os.path.join(src, s["Mesh"][0]), os.path.join(out, s["Mesh"][0]))
    for b in bsdfs.values():
        img = b["Principled"]["color"].get("Image")
        if img:
            rel = img.replace("\\", "/")
            shutil.copyfile(os.path.join(src, rel), os.path.join(out, rel))
    shutil.copyfile(os.path.join(src, "living_room.mtl"), os.path.join(out, "living_room.mtl"))
    scene = {"named_bsdfs": bsdfs, "camera": room["camera"], "lights": room["lights"], "shapes": shapes}
    with open(os.path.join(out, "scene.json"), "w") as f:
        json.dump(scene, f, indent=4)
    print(f"{len(shapes)} shapes, {len(bsdfs)} BSDFs -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/scenes")
