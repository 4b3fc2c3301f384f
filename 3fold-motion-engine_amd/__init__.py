"""MI355X-native contact step (HIP/gfx950) behind the reference ContactEngine protocol."""
