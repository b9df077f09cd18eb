"""MI355X-native implicit soft-body FEM step (SuperDex Physics soft path)."""
