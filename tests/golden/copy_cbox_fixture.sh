#!/bin/sh
# Copies the reference's shipped Cornell-box scene data (Mitsuba XML + OBJ
# meshes; data, not code) into tests/golden/cbox/ as loader and end-to-end
# fixtures.  Parenthesised file names are made shell-friendly; the XML text is
# otherwise unchanged.  Run from the repository root in a container that has
# the reference checkout at /root/reference.
set -eu
src=/root/reference/scene/CBox
dst=tests/golden/cbox
mkdir -p "$dst/meshes"
cp "$src/cbox.xml" "$dst/cbox.xml"
cp "$src/cbox(mesh).xml" "$dst/cbox_mesh.xml"
cp "$src/cbox(sphere).xml" "$dst/cbox_sphere.xml"
for m in backwall leftwall light rightwall sphere1 sphere2 walls; do
    cp "$src/meshes/$m.obj" "$dst/meshes/$m.obj"
done
# The reference's published render of cbox(sphere).xml (800x600, 3006 samples,
# images/3006spp.png): the converged-image fixture of the end-to-end test.
cp /root/reference/images/3006spp.png "$dst/cbox_sphere_3006spp.png"
