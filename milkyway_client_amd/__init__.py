"""MI355X-native separation likelihood (milkyway_separation) behind the
reference evaluator's C interface."""
