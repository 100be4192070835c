#!/bin/sh
# Copies the Forrestania data files of the reference checkout into this folder, as test
# fixtures: the eight unit surfaces and the label grid of the baseline model, the unit
# property table, the gravity stations with the magnetic data sampled at them, and the
# two run configuration files of the Forrestania app.
# Data only; nothing under the reference is run.
set -e
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
BASE=$REF/apps/forrestania/User_Perturbed_Physical_Properties/baseline
for k in 1 2 3 4 5 6 7 8; do
  cp "$BASE/meshes/volume_group_$k.ts" "$HERE/"
done
cp "$BASE/meshes/label_grid.bin" "$HERE/"
cp "$BASE/cluster_properties.csv" "$HERE/"
cp "$REF/datasets/Forrestania/Forrestania_Gravity_Station_trim_.csv" "$HERE/"
cp "$REF/datasets/Forrestania/Forrestania_mag_at_gravity_stations.csv" "$HERE/"
cp "$REF/apps/forrestania/test_config.ini" "$REF/apps/forrestania/test_config_real.ini" "$HERE/"
chmod u+w "$HERE"/*.ts "$HERE"/*.bin "$HERE"/*.csv "$HERE"/*.ini
