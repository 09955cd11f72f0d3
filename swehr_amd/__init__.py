"""MI355X-native SWEHR coupled time step (package stub; filled in below)."""
