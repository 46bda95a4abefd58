#!/bin/sh
# Copies the data files used as fixtures from the reference checkout.
set -e
cp /root/reference/config/maos/bin/NapGaussian.bin "$(dirname "$0")/NapGaussian.bin"
