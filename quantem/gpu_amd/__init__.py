"""MI355X-native single-sideband ptychography backend for quantem."""
