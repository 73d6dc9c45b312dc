"""Input and output formats on either side of the batch step: reaction
criteria (rxns.xml) and results.json."""
