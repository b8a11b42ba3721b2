"""The serving precedent's JSON documents on the CPU: the query document of
README.md:115-133 and the configuration document, validated as
src/dist_genie/parser.cu:16-217 validates them (each refusal carries the
reference's message)."""
import json

import pytest

from genie_amd import GenieError, wire

README_QUERY = """{
  "topk": 10,
  "queries": [
    {
      "content": [1, 2, 3, 4, 5],
      "clusters": [0, 9, 13]
    },
    {
      "content": [90, 24, 33, 14, 5],
      "clusters": [3, 7]
    }
  ]
}"""

CONFIG = dict(data_file="sift", dim=5, count_threshold=5, data_type=0, search_type=0, max_data_size=0,
              num_of_cluster=20, data_format=0, radius=1, selectivity=0.0)


def test_readme_query_document():
    assert wire.parse_query(README_QUERY) == (10, 2)


@pytest.mark.parametrize("doc,msg", [
    ('{"queries": []}', "Entry 'topk' is missing"),
    ('{"topk": "10", "queries": []}', "Entry 'topk' should be an interger"),
    ('{"topk": 10.0, "queries": []}', "Entry 'topk' should be an interger"),
    ('{"topk": 10}', "Entry 'queries' is missing"),
    ('{"topk": 10, "queries": {}}', "Entry 'queries' should be an array"),
    ('{"topk": 10, "queries": [{"clusters": [0]}]}', "Some query misses the 'content' section"),
    ('{"topk": 10, "queries": [{"content": 5, "clusters": [0]}]}', "Query's 'content' section should be an array"),
    ('{"topk": 10, "queries": [{"content": [1]}]}', "Some query misses the 'clusters' section"),
    ('{"topk": 10, "queries": [{"content": [1], "clusters": 0}]}', "Query's 'clusters' section should be an array"),
    ('{"topk": 10, "queries": [{"content": [1.5], "clusters": [0]}]}', "'content' values should be integers"),
    ('{"topk": 10, "queries": [],}', "not a valid JSON document"),
    ('{"topk": 10 "queries": []}', "not a valid JSON document"),
    ('[1, 2]', "not a JSON object"),
    ('{"topk": 10, "queries": []} x', "not a valid JSON document"),
])
def test_query_document_refusals(doc, msg):
    with pytest.raises(GenieError) as e:
        wire.parse_query(doc)
    assert msg in str(e.value)


def test_json_reader_accepts_the_grammar():
    doc = ' {"topk" : 3 ,"queries":[ {"clusters":[],"content":[-1,0,2147483647],"note":"\\u00e9\\n\\"x\\"",' \
          '"nested":{"a":[true,false,null,1e3,-0.5]}} ] } \n'
    assert wire.parse_query(doc) == (3, 1)
    with pytest.raises(GenieError):
        wire.parse_query('{"topk": 3, "queries": [{"content": [2147483648], "clusters": []}]}')


def test_configuration_document():
    c = wire.parse_config(json.dumps(CONFIG))
    assert c == dict(dim=5, radius=1, num_of_cluster=20, data_format=0, selectivity=0.0)
    c = wire.parse_config(json.dumps(dict(CONFIG, selectivity=0.25)))
    assert abs(c["selectivity"] - 0.25) < 1e-7


@pytest.mark.parametrize("change,msg", [
    (dict(data_file=None), "Entry data_file is missing"),
    (dict(data_file=3), "Entry data_file has wrong type"),
    (dict(dim=None), "Entry dim is missing"),
    (dict(dim="5"), "Entry dim has wrong type"),
    (dict(radius=1.0), "Entry radius has wrong type"),
    (dict(selectivity=None), "Entry selectivity is missing"),
    (dict(selectivity=0), "Entry selectivity has wrong type"),  # an integer literal is not a float (rapidjson)
])
def test_configuration_refusals(change, msg):
    cfg = dict(CONFIG)
    for k, v in change.items():
        if v is None:
            del cfg[k]
        else:
            cfg[k] = v
    with pytest.raises(GenieError) as e:
        wire.parse_config(json.dumps(cfg))
    assert msg in str(e.value)
