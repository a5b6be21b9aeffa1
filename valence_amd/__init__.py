"""MI355X-native VSVB energy path (VALENCE drop-in for the energy call)."""
